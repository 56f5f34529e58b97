// serialize.hip -- saving and loading an assembled operator.
//
// The file holds what the cross approximations produced -- per family, level and
// box the rank, the row and column pivots (tree-order point indices), L and R and
// the per-step history; per semi-nested / non-nested pair the same -- together
// with the input points and the build parameters.  Loading rebuilds the plan from
// the points, puts the skeletons back and runs the operator phase of the assembly
// (transfer, M2L, near-field and CUR blocks, FP32 copies, work lists, inner
// operator) with the same kernels, so the loaded operator is the saved one to the
// last bit while no cross approximation runs.  Operator pools are not stored:
// re-evaluating them costs a fraction of the assembly and keeps the file at the
// size of the skeletons.
//
// Layout (little-endian): "H2WSKEL1", int32 dim, variant, choice, tol_pow, n_max,
// nlevels, compat, 0; int64 N; double L; dim x N doubles (input order); per
// family f = 0, 1 and level j = 0..nlevels: int32 nb, then per box int32 rank,
// rank x int32 Ir, rank x int32 Ic, 2 rank^2 doubles (L then R), int32 nhist,
// nhist doubles; int64 npairs, per pair int32 level, b, v, rank, Ir, Ic, L R;
// uint64 FNV-1a hash of everything before it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <array>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/h2w.h"
#include "h2w_handle.hpp"

namespace h2w {
namespace {

const char kMagic[9] = "H2WSKEL1";

struct Writer {
    std::FILE *f;
    uint64_t hash = 1469598103934665603ULL;
    void raw(const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < n; ++i) hash = (hash ^ c[i]) * 1099511628211ULL;
        if (n && std::fwrite(p, 1, n, f) != n) throw Error{H2W_EINVAL, "write failed"};
    }
    template <typename T>
    void put(T v) { raw(&v, sizeof(T)); }
    template <typename T>
    void vec(const std::vector<T> &v, size_t n) {
        if (v.size() < n) throw Error{H2W_EINVAL, "inconsistent skeleton"};
        raw(v.data(), n * sizeof(T));
    }
};

struct Reader {
    std::FILE *f;
    uint64_t hash = 1469598103934665603ULL;
    void raw(void *p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) throw Error{H2W_EINVAL, "truncated operator file"};
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < n; ++i) hash = (hash ^ c[i]) * 1099511628211ULL;
    }
    template <typename T>
    T get() {
        T v;
        raw(&v, sizeof(T));
        return v;
    }
    template <typename T>
    void vec(std::vector<T> &v, int64_t n, int64_t cap) {
        if (n < 0 || n > cap) throw Error{H2W_EINVAL, "corrupt operator file (size out of range)"};
        v.resize((size_t)n);
        raw(v.data(), (size_t)n * sizeof(T));
    }
};

struct FileCloser {
    void operator()(std::FILE *f) const {
        if (f) std::fclose(f);
    }
};

// row set of every box of a family: the leaf's points, else the children's row pivots
void derive_S(Handle &h, int f, int lo) {
    const Plan &P = h.plan;
    auto &F = h.fam[f];
    for (int j = P.nL; j >= lo; --j) {
        FamLevel &fl = F[j];
        fl.S.assign(fl.nb, {});
        for (int b = 0; b < fl.nb; ++b) {
            if (j == P.nL) {
                for (int64_t t = 0; t < P.bcount[j][b]; ++t) fl.S[b].push_back((int32_t)(P.bstart[j][b] + t));
            } else {
                for (int c = 0; c < P.nchild; ++c) {
                    const auto &ir = F[j + 1].Ir[b * P.nchild + c];
                    fl.S[b].insert(fl.S[b].end(), ir.begin(), ir.end());
                }
            }
        }
    }
}

}  // namespace
}  // namespace h2w

using h2w::Error;

extern "C" {

int h2w_save(const h2w_handle *h, const char *path) {
    try {
        if (!h || !path) throw Error{H2W_EINVAL, "NULL argument"};
        if (!h->assembled) throw Error{H2W_ESTATE, "h2w_save before h2w_assemble"};
        if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: operator files hold real skeleton factors only"};
        std::unique_ptr<std::FILE, h2w::FileCloser> fp(std::fopen(path, "wb"));
        if (!fp) throw Error{H2W_EINVAL, std::string("cannot open ") + path};
        h2w::Writer w{fp.get()};
        const h2w::Plan &P = h->plan;
        w.raw(h2w::kMagic, 8);
        const int32_t hdr[8] = {h->dim, h->variant, h->choice, h->tol_pow, h->n_max, P.nL, h->opts.compat, 0};
        w.raw(hdr, sizeof(hdr));
        w.put<int64_t>(h->N);
        w.put<double>(h->L);
        // input-order points
        const std::vector<double> *tc[3] = {&P.px, &P.py, &P.pz};
        std::vector<double> in(h->N);
        for (int d = 0; d < h->dim; ++d) {
            for (int64_t t = 0; t < h->N; ++t) in[P.perm[t]] = (*tc[d])[t];
            w.vec(in, h->N);
        }
        for (int f = 0; f < 2; ++f)
            for (int j = 0; j <= P.nL; ++j) {
                const bool have = j < (int)h->fam[f].size();
                const h2w::FamLevel *fl = have ? &h->fam[f][j] : nullptr;
                const int32_t nb = fl ? (int32_t)fl->rank.size() : 0;
                w.put<int32_t>(nb);
                for (int b = 0; b < nb; ++b) {
                    const int32_t r = fl->rank[b];
                    w.put<int32_t>(r);
                    w.vec(fl->Ir[b], r);
                    w.vec(fl->Ic[b], r);
                    w.vec(fl->lr[b], (size_t)2 * r * r);
                    w.put<int32_t>((int32_t)fl->hist[b].size());
                    w.vec(fl->hist[b], fl->hist[b].size());
                }
            }
        w.put<int64_t>((int64_t)h->pairs.size());
        for (const auto &p : h->pairs) {
            const int32_t v[4] = {p.level, p.b, p.v, p.rank};
            w.raw(v, sizeof(v));
            w.vec(p.Ir, p.rank);
            w.vec(p.Ic, p.rank);
            w.vec(p.lr, (size_t)2 * p.rank * p.rank);
        }
        const uint64_t hash = w.hash;
        w.put<uint64_t>(hash);
        if (std::fflush(fp.get()) != 0) throw Error{H2W_EINVAL, "write failed"};
        return H2W_OK;
    } catch (const Error &e) {
        h2w::set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        h2w::set_error("host allocation failed");
        return H2W_ENOMEM;
    }
}

int h2w_load(const char *path, const h2w_opts *opts, h2w_handle **out) {
    h2w_handle *h = nullptr;
    try {
        if (!path || !out) throw Error{H2W_EINVAL, "NULL argument"};
        *out = nullptr;
        std::unique_ptr<std::FILE, h2w::FileCloser> fp(std::fopen(path, "rb"));
        if (!fp) throw Error{H2W_EINVAL, std::string("cannot open ") + path};
        h2w::Reader r{fp.get()};
        char magic[8];
        r.raw(magic, 8);
        if (std::memcmp(magic, h2w::kMagic, 8) != 0) throw Error{H2W_EINVAL, "not an h2w operator file"};
        int32_t hdr[8];
        r.raw(hdr, sizeof(hdr));
        const int dim = hdr[0], variant = hdr[1], choice = hdr[2], tol_pow = hdr[3], n_max = hdr[4], nL = hdr[5];
        const int64_t N = r.get<int64_t>();
        const double L = r.get<double>();
        if (dim < 1 || dim > 3 || nL < 1 || nL > 20 || N < 1 || N > ((int64_t)1 << 31))
            throw Error{H2W_EINVAL, "corrupt operator file (header)"};
        std::vector<double> pts, tmp;
        for (int d = 0; d < dim; ++d) {
            r.vec(tmp, N, N);
            pts.insert(pts.end(), tmp.begin(), tmp.end());
        }
        h2w_opts o;
        h2w_default_opts(&o);
        if (opts) o = *opts;
        o.nlevels = nL;
        o.compat = hdr[6];
        // The file is read and checked completely against a host-only plan before any device
        // is touched: every index it holds is used to address points or pool blocks later.
        const int want_device = !o.plan_only;
        o.plan_only = 1;
        int rc = h2w_create(dim, variant, N, pts.data(), n_max, L, tol_pow, choice, &o, &h);
        if (rc != H2W_OK) throw Error{rc, h2w_last_error()};
        const h2w::Plan &P = h->plan;
        const double t0 = h2w::now_s();
        auto pivots_ok = [&](const std::vector<int32_t> &v) {
            for (int32_t i : v)
                if (i < 0 || i >= N) return false;
            return true;
        };
        // levels each family's skeletons exist on (far: 2..nL, vertex: 1..nL; see the assembly)
        const bool fam_used[2] = {dim >= 2 && variant != H2W_NON_NESTED, variant == H2W_NESTED};
        const int fam_lo[2] = {2, 1};
        for (int f = 0; f < 2; ++f) {
            h->fam[f].assign(P.nL + 1, h2w::FamLevel());
            for (int j = 0; j <= P.nL; ++j) {
                const int32_t nb = r.get<int32_t>();
                const bool used = fam_used[f] && j >= fam_lo[f];
                if (nb != (used ? P.nboxes[j] : 0)) throw Error{H2W_EINVAL, "corrupt operator file (box count)"};
                h2w::FamLevel &fl = h->fam[f][j];
                fl.nb = nb;
                fl.rank.resize(nb);
                fl.Ir.resize(nb);
                fl.Ic.resize(nb);
                fl.lr.resize(nb);
                fl.hist.resize(nb);
                for (int b = 0; b < nb; ++b) {
                    const int32_t rk = r.get<int32_t>();
                    if (rk < 0 || rk > P.bcount[j][b]) throw Error{H2W_EINVAL, "corrupt operator file (rank)"};
                    fl.rank[b] = rk;
                    r.vec(fl.Ir[b], rk, N);
                    r.vec(fl.Ic[b], rk, N);
                    r.vec(fl.lr[b], (int64_t)2 * rk * rk, (int64_t)2 * rk * rk);
                    const int32_t nh = r.get<int32_t>();
                    r.vec(fl.hist[b], nh, (int64_t)2 * N);
                    if (!pivots_ok(fl.Ir[b]) || !pivots_ok(fl.Ic[b]))
                        throw Error{H2W_EINVAL, "corrupt operator file (pivot)"};
                }
            }
        }
        // a far skeleton is taken from its row set: the leaf's points, else the children's skeletons
        if (fam_used[0])
            for (int j = P.nL - 1; j >= fam_lo[0]; --j)
                for (int b = 0; b < P.nboxes[j]; ++b) {
                    int64_t rows = 0;
                    for (int c = 0; c < P.nchild; ++c) rows += h->fam[0][j + 1].rank[b * P.nchild + c];
                    if (h->fam[0][j].rank[b] > rows) throw Error{H2W_EINVAL, "corrupt operator file (rank)"};
                }
        // pairs: exactly the directed pairs the assembly makes, in its order
        std::vector<std::array<int, 3>> expect;
        if (variant != H2W_NESTED)
            for (int j = 1; j <= P.nL; ++j)
                for (int b = 0; b < P.nboxes[j]; ++b)
                    for (int v : (variant == H2W_NON_NESTED ? P.il[j][b] : P.vert_of(j, b))) expect.push_back({j, b, v});
        const int64_t np = r.get<int64_t>();
        if (np != (int64_t)expect.size()) throw Error{H2W_EINVAL, "corrupt operator file (pair count)"};
        h->pairs.assign((size_t)np, h2w::Pair());
        for (int64_t q = 0; q < np; ++q) {
            h2w::Pair &p = h->pairs[q];
            int32_t v[4];
            r.raw(v, sizeof(v));
            if (v[0] != expect[q][0] || v[1] != expect[q][1] || v[2] != expect[q][2] || v[3] < 0 ||
                v[3] > std::min(P.bcount[v[0]][v[1]], P.bcount[v[0]][v[2]]))
                throw Error{H2W_EINVAL, "corrupt operator file (pair)"};
            p.level = v[0];
            p.b = v[1];
            p.v = v[2];
            p.rank = v[3];
            r.vec(p.Ir, p.rank, N);
            r.vec(p.Ic, p.rank, N);
            r.vec(p.lr, (int64_t)2 * p.rank * p.rank, (int64_t)2 * p.rank * p.rank);
            if (!pivots_ok(p.Ir) || !pivots_ok(p.Ic)) throw Error{H2W_EINVAL, "corrupt operator file (pair pivot)"};
        }
        const uint64_t want = r.hash;
        const uint64_t got = r.get<uint64_t>();
        if (got != want) throw Error{H2W_EINVAL, "operator file checksum mismatch"};
        char extra;
        if (std::fread(&extra, 1, 1, fp.get()) != 0) throw Error{H2W_EINVAL, "corrupt operator file (trailing bytes)"};
        if (P.dim >= 2 && variant != H2W_NON_NESTED) h2w::derive_S(*h, 0, 2);
        if (variant == H2W_NESTED) h2w::derive_S(*h, 1, 1);
        if (!want_device) {  // checked only: no device work
            *out = h;
            return H2W_OK;
        }
        h2w::attach_device(*h);
        h2w::rebuild_operators(*h);
        h->assembly_seconds = h2w::now_s() - t0;
        *out = h;
        return H2W_OK;
    } catch (const Error &e) {
        h2w::set_error(e.msg);
        if (h) h2w_destroy(h);
        return e.code;
    } catch (const std::bad_alloc &) {
        h2w::set_error("host allocation failed");
        if (h) h2w_destroy(h);
        return H2W_ENOMEM;
    }
}

}  // extern "C"
