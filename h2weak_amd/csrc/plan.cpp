// plan.cpp -- tree, tree-order permutation and interaction lists (host).
//
// Behaviour follows the reference's tree set-up:
//   child numbering / centres ...... 3D/codes/HODLR3Dtree.hpp:1700-1741, 2D/codes/HODLR2Dtree.hpp:426-449,
//                                    1D/codes/HODLR1Dtree.hpp:150-165
//   point bucketing ("<= centre") .. 3D/codes/HODLR3Dtree.hpp:1769-1837 (2D :478-518, 1D :168-196)
//   tree order (children concat) ... 3D/codes/HODLR3Dtree.hpp:1839-1852
//   near / interaction lists ....... 3D/codes/HODLR3Dtree.hpp:155-1680, 2D/codes/HODLR2Dtree.hpp:142-423,
//                                    1D/codes/HODLR1Dtree.hpp:116-147
//   vertex flag .................... 3D/codes/HODLR3Dtree.hpp:1744-1767
//
// The reference encodes the lists in hand-written per-child tables.  Here they
// come from integer box coordinates: near(B) = boxes at offset d with
// max|d_i| <= 1 sharing a face or an edge (1 <= #nonzero(d) <= dim-1);
// IL(B) = the vertex-sharing siblings, then the children of the parent's near
// neighbours (taken in slot order) that are not near(B).  In 3D the table order
// lists a neighbour's vertex-sharing child first; that order is kept.  A
// vertex-sharing box is one with |d_i| = 1 in every coordinate, which is what
// the reference's centre-distance test (<= 2 sqrt(dim) r_j) selects.
#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>

#include "h2w_internal.hpp"

namespace h2w {

namespace {

void child_bits(int dim, int c, int *b) {
    b[0] = b[1] = b[2] = 0;
    if (dim == 1) {
        b[0] = c;
    } else if (dim == 2) {
        b[0] = (c == 1 || c == 2);
        b[1] = (c == 2 || c == 3);
    } else {
        int cc = c & 3;
        b[0] = (cc == 1 || cc == 2);
        b[1] = (cc == 2 || cc == 3);
        b[2] = (c >= 4);
    }
}

// neighbour offsets in the reference's slot order
std::vector<std::array<int, 3>> near_offsets(int dim) {
    std::vector<std::array<int, 3>> v;
    if (dim == 2) {
        v.push_back({0, -1, 0});
        v.push_back({1, 0, 0});
        v.push_back({0, 1, 0});
        v.push_back({-1, 0, 0});
    } else if (dim == 3) {
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int nz = (dx != 0) + (dy != 0) + (dz != 0);
                    if (nz == 0 || nz == 3) continue;
                    v.push_back({dx, dy, dz});
                }
    }
    return v;
}

}  // namespace

std::vector<int32_t> Plan::vert_of(int j, int b) const {
    std::vector<int32_t> v;
    for (size_t i = 0; i < il[j][b].size(); ++i)
        if (ilv[j][b][i]) v.push_back(il[j][b][i]);
    return v;
}
std::vector<int32_t> Plan::far_of(int j, int b) const {
    std::vector<int32_t> v;
    for (size_t i = 0; i < il[j][b].size(); ++i)
        if (!ilv[j][b][i]) v.push_back(il[j][b][i]);
    return v;
}

void Plan::build(int dim_, int64_t N_, const double *pts, double L_, int nlevels) {
    dim = dim_;
    N = N_;
    L = L_;
    nL = nlevels;
    nchild = 1 << dim;
    nboxes.assign(nL + 1, 1);
    for (int j = 1; j <= nL; ++j) nboxes[j] = nboxes[j - 1] * nchild;
    std::vector<double> radius(nL + 1, L);
    for (int j = 1; j <= nL; ++j) radius[j] = 0.5 * radius[j - 1];

    // leaf box of every point by descending the centres (same comparisons and
    // same centre arithmetic as the level-by-level bucketing of the reference)
    const double *X = pts, *Y = dim > 1 ? pts + N : nullptr, *Z = dim > 2 ? pts + 2 * N : nullptr;
    std::vector<int32_t> leaf(N);
    for (int64_t i = 0; i < N; ++i) {
        double c[3] = {0.0, 0.0, 0.0};
        int64_t k = 0;
        double p[3] = {X[i], Y ? Y[i] : 0.0, Z ? Z[i] : 0.0};
        for (int j = 0; j < nL; ++j) {
            double shift = 0.5 * radius[j];
            bool lo[3];
            for (int d = 0; d < dim; ++d) lo[d] = p[d] <= c[d];
            int ch;
            if (dim == 1) ch = lo[0] ? 0 : 1;
            else if (dim == 2) ch = lo[0] ? (lo[1] ? 0 : 3) : (lo[1] ? 1 : 2);
            else {
                ch = lo[0] ? (lo[1] ? 0 : 3) : (lo[1] ? 1 : 2);
                if (!lo[2]) ch += 4;
            }
            int b[3];
            child_bits(dim, ch, b);
            for (int d = 0; d < dim; ++d) c[d] = b[d] ? c[d] + shift : c[d] - shift;
            k = k * nchild + ch;
        }
        leaf[i] = (int32_t)k;
    }
    // stable counting sort by leaf: within a leaf the input order is kept
    int nleaf = nboxes[nL];
    std::vector<int64_t> cnt(nleaf + 1, 0);
    for (int64_t i = 0; i < N; ++i) cnt[leaf[i] + 1]++;
    for (int b = 0; b < nleaf; ++b) cnt[b + 1] += cnt[b];
    perm.assign(N, 0);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t i = 0; i < N; ++i) perm[pos[leaf[i]]++] = (int32_t)i;
    }
    px.assign(N, 0.0);
    py.assign(N, 0.0);
    pz.assign(N, 0.0);
    for (int64_t t = 0; t < N; ++t) {
        px[t] = X[perm[t]];
        if (Y) py[t] = Y[perm[t]];
        if (Z) pz[t] = Z[perm[t]];
    }
    bstart.assign(nL + 1, {});
    bcount.assign(nL + 1, {});
    bstart[nL].resize(nleaf);
    bcount[nL].resize(nleaf);
    for (int b = 0; b < nleaf; ++b) {
        bstart[nL][b] = cnt[b];
        bcount[nL][b] = cnt[b + 1] - cnt[b];
    }
    for (int j = nL - 1; j >= 0; --j) {
        bstart[j].resize(nboxes[j]);
        bcount[j].resize(nboxes[j]);
        for (int b = 0; b < nboxes[j]; ++b) {
            bstart[j][b] = bstart[j + 1][b * nchild];
            int64_t s = 0;
            for (int c = 0; c < nchild; ++c) s += bcount[j + 1][b * nchild + c];
            bcount[j][b] = s;
        }
    }

    // integer coordinates and coordinate -> box maps
    std::vector<std::vector<std::array<int, 3>>> ic(nL + 1);
    std::vector<std::vector<int32_t>> at(nL + 1);
    ic[0].push_back({0, 0, 0});
    at[0].push_back(0);
    for (int j = 1; j <= nL; ++j) {
        ic[j].resize(nboxes[j]);
        int side = 1 << j;
        at[j].assign((size_t)nboxes[j], -1);
        for (int b = 0; b < nboxes[j]; ++b) {
            int p = b / nchild, c = b % nchild;
            int bits[3];
            child_bits(dim, c, bits);
            for (int d = 0; d < 3; ++d) ic[j][b][d] = d < dim ? 2 * ic[j - 1][p][d] + bits[d] : 0;
            int64_t lin = 0;
            for (int d = dim - 1; d >= 0; --d) lin = lin * side + ic[j][b][d];
            at[j][lin] = b;
        }
    }
    icoord = ic;
    // bounding boxes of the boxes' points, leaves up
    bbox.assign(nL + 1, {});
    for (int j = nL; j >= 0; --j) {
        bbox[j].assign(nboxes[j], {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f});
        for (int b = 0; b < nboxes[j]; ++b) {
            auto &bb = bbox[j][b];
            if (j == nL) {
                for (int64_t t = bstart[j][b]; t < bstart[j][b] + bcount[j][b]; ++t) {
                    const double c[3] = {px[t], dim > 1 ? py[t] : 0.0, dim > 2 ? pz[t] : 0.0};
                    for (int d = 0; d < 3; ++d) {
                        bb[d] = std::min(bb[d], (float)c[d]);
                        bb[3 + d] = std::max(bb[3 + d], (float)c[d]);
                    }
                }
            } else {
                for (int c = 0; c < nchild; ++c) {
                    const auto &cb = bbox[j + 1][b * nchild + c];
                    for (int d = 0; d < 3; ++d) {
                        bb[d] = std::min(bb[d], cb[d]);
                        bb[3 + d] = std::max(bb[3 + d], cb[3 + d]);
                    }
                }
            }
        }
    }
    auto lookup = [&](int j, const std::array<int, 3> &q) -> int {
        int side = 1 << j;
        int64_t lin = 0;
        for (int d = dim - 1; d >= 0; --d) {
            if (q[d] < 0 || q[d] >= side) return -1;
            lin = lin * side + q[d];
        }
        return at[j][lin];
    };
    auto offs = near_offsets(dim);
    auto near_list = [&](int j, int b) {
        std::vector<int32_t> v;
        for (auto &o : offs) {
            std::array<int, 3> q = {ic[j][b][0] + o[0], ic[j][b][1] + o[1], ic[j][b][2] + o[2]};
            int n = lookup(j, q);
            if (n >= 0) v.push_back(n);
        }
        return v;
    };
    auto rel = [&](int j, int a, int b, int *d) {
        int mx = 0, nz = 0;
        for (int q = 0; q < dim; ++q) {
            d[q] = ic[j][b][q] - ic[j][a][q];
            mx = std::max(mx, std::abs(d[q]));
            nz += d[q] != 0;
        }
        // 0: far, 1: near (face/edge), 2: vertex, 3: self
        if (nz == 0) return 3;
        if (mx > 1) return 0;
        return nz == dim ? 2 : 1;
    };

    il.assign(nL + 1, {});
    ilv.assign(nL + 1, {});
    for (int j = 1; j <= nL; ++j) {
        il[j].assign(nboxes[j], {});
        ilv[j].assign(nboxes[j], {});
        for (int b = 0; b < nboxes[j]; ++b) {
            int p = b / nchild;
            auto &lst = il[j][b];
            auto &vf = ilv[j][b];
            int d[3];
            for (int c2 = 0; c2 < nchild; ++c2) {
                int s = p * nchild + c2;
                if (s == b) continue;
                if (rel(j, b, s, d) == 2) {
                    lst.push_back(s);
                    vf.push_back(1);
                }
            }
            if (j >= 2) {
                for (int pn : near_list(j - 1, p)) {
                    std::vector<int> order;
                    int vc = -1;
                    if (dim == 3)
                        for (int c2 = 0; c2 < nchild; ++c2)
                            if (rel(j, b, pn * nchild + c2, d) == 2) vc = c2;
                    if (vc >= 0) order.push_back(vc);
                    for (int c2 = 0; c2 < nchild; ++c2)
                        if (c2 != vc) order.push_back(c2);
                    for (int c2 : order) {
                        int s = pn * nchild + c2;
                        int r = rel(j, b, s, d);
                        if (r == 1 || r == 3) continue;
                        lst.push_back(s);
                        vf.push_back(r == 2 ? 1 : 0);
                    }
                }
            }
        }
    }
    near.assign(nboxes[nL], {});
    for (int b = 0; b < nboxes[nL]; ++b) near[b] = near_list(nL, b);
}

}  // namespace h2w
