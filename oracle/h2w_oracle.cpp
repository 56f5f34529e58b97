// h2w_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU FP64 restatement of the RiteshKhan/H2weak hot path, used as the
// checker for the HIP library (tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may call it; nothing in h2weak_amd/ may).
// It is written from the reference's behaviour, function by function,
// without Eigen:
//
//   kernel entries + diagonal rules ... */codes/kernel.hpp (1D:95-185, 2D:101-318, 3D:101-192)
//   GL-16 diagonal quadrature ......... */codes/integral{1,2,3}d.hpp; nodes by Newton on P16,
//                                       pinned against Gauss_Legendre_Nodes_and_Weights.hpp:471-523
//   partial-pivot ACA (3 variants) .... 2D/codes/aca.hpp:23-38,158-294,296-446; 1D/codes/aca.hpp:40-156
//   tree, centres, neighbour slots,
//   interaction lists, vertex flags ... 2D/codes/HODLR2Dtree.hpp:104-534, 3D/codes/HODLR3Dtree.hpp:116-1852,
//                                       1D/codes/HODLR1Dtree.hpp:79-212
//   far NCA bottom-up + L2P ........... 3D/codes/HODLR3Dtree.hpp:1872-1939
//   vertex NCA top-down + M2M_ver ..... 3D/codes/HODLR3Dtree.hpp:2062-2133; 1D/codes/HODLR1Dtree.hpp:231-302
//   M2L far/ver ....................... 3D/codes/HODLR3Dtree.hpp:1978-1998,2164-2185
//   semi-nested CUR + apply ........... 3D/codes/HODLR3Dtree.hpp:2250-2351; 1D UV 1D/codes/HODLR1Dtree.hpp:418-463
//   near field ........................ 3D/codes/HODLR3Dtree.hpp:2375-2423
//   matvec passes + collect/reorder ... 3D/codes/HODLR3Dtree.hpp:1941-2055,2135-2237,2426-2503
//   storage count ..................... 3D/codes/HODLR3Dtree.hpp:2505-2559,2585-2628; 1D :535-588; 2D :1207-1305
//   dense product ..................... 3D/codes/HODLR3Dtree.hpp:2630-2643
//   GMRES (MGS, Givens, restart=max) .. 1D/codes/gmres.hpp:30-200
//
// Parallel structure follows the reference: OpenMP over boxes inside a level
// for every evaluate pass, over leaves for the near-field assembly, and inside
// each row/column gather of the ACA; the per-box assembly loops are serial.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <omp.h>
#include <vector>

namespace {

const double PI = 4.0 * std::atan(1.0);

// ---------------------------------------------------------------- matrices
// Column-major dense matrix (the reference's Eigen::MatrixXd layout).
struct Mat {
    int r = 0, c = 0;
    std::vector<double> a;
    Mat() {}
    Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
    double &operator()(int i, int j) { return a[(size_t)j * r + i]; }
    double operator()(int i, int j) const { return a[(size_t)j * r + i]; }
    size_t size() const { return (size_t)r * c; }
};
typedef std::vector<double> Vec;

Mat transpose(const Mat &A) {
    Mat T(A.c, A.r);
    for (int j = 0; j < A.c; ++j)
        for (int i = 0; i < A.r; ++i) T(j, i) = A(i, j);
    return T;
}
// y (+)= A x
void gemv(const Mat &A, const double *x, double *y, bool acc) {
    if (!acc) std::fill(y, y + A.r, 0.0);
    for (int j = 0; j < A.c; ++j) {
        double xj = x[j];
        for (int i = 0; i < A.r; ++i) y[i] += A(i, j) * xj;
    }
}
// y (+)= A^T x
void gemv_t(const Mat &A, const double *x, double *y, bool acc) {
    for (int j = 0; j < A.c; ++j) {
        double s = 0.0;
        for (int i = 0; i < A.r; ++i) s += A(i, j) * x[i];
        y[j] = acc ? y[j] + s : s;
    }
}
// X R = A with R upper (Eigen triangularView<Upper>().solve<OnTheRight>)
Mat solve_right_upper(const Mat &R, const Mat &A) {
    Mat X(A.r, A.c);
    for (int i = 0; i < A.r; ++i)
        for (int k = 0; k < A.c; ++k) {
            double s = A(i, k);
            for (int l = 0; l < k; ++l) s -= X(i, l) * R(l, k);
            X(i, k) = s / R(k, k);
        }
    return X;
}
// Y L = A with L lower (Eigen triangularView<Lower>().solve<OnTheRight>)
Mat solve_right_lower(const Mat &L, const Mat &A) {
    Mat Y(A.r, A.c);
    for (int i = 0; i < A.r; ++i)
        for (int k = A.c - 1; k >= 0; --k) {
            double s = A(i, k);
            for (int l = k + 1; l < A.c; ++l) s -= Y(i, l) * L(l, k);
            Y(i, k) = s / L(k, k);
        }
    return Y;
}
// (L R)^{-T} Z : R^T W = Z (lower), then L^T Y = W (upper)
Mat solve_lr_transpose(const Mat &L, const Mat &R, const Mat &Z) {
    int r = L.r;
    Mat W(Z.r, Z.c), Y(Z.r, Z.c);
    for (int c = 0; c < Z.c; ++c) {
        for (int k = 0; k < r; ++k) {
            double s = Z(k, c);
            for (int l = 0; l < k; ++l) s -= R(l, k) * W(l, c);
            W(k, c) = s / R(k, k);
        }
        for (int k = r - 1; k >= 0; --k) {
            double s = W(k, c);
            for (int l = k + 1; l < r; ++l) s -= L(l, k) * Y(l, c);
            Y(k, c) = s / L(k, k);
        }
    }
    return Y;
}
// t = R^{-1} L^{-1} v (in place)
void solve_lr_vec(const Mat &L, const Mat &R, double *v) {
    int r = L.r;
    for (int k = 0; k < r; ++k) {
        double s = v[k];
        for (int l = 0; l < k; ++l) s -= L(k, l) * v[l];
        v[k] = s / L(k, k);
    }
    for (int k = r - 1; k >= 0; --k) {
        double s = v[k];
        for (int l = k + 1; l < r; ++l) s -= R(k, l) * v[l];
        v[k] = s / R(k, k);
    }
}
double vnorm(const Vec &v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}
double vdot(const Vec &a, const Vec &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// ---------------------------------------------------------------- GL-16
// Gauss-Legendre nodes by Newton iteration on P_16, ordered as the
// reference table orders them (pairs -x,+x by increasing |x|).
void gauss_legendre16(double *nodes, double *weights) {
    const int n = 16;
    double pos[8], wts[8];
    for (int i = 0; i < n / 2; ++i) {
        double x = std::cos(PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-17) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (x * p1 - p0) / (x * x - 1.0);
        pos[i] = x;  // decreasing
        wts[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    for (int i = 0; i < n / 2; ++i) {
        int s = n / 2 - 1 - i;  // increasing |x|
        nodes[2 * i] = -pos[s];
        nodes[2 * i + 1] = pos[s];
        weights[2 * i] = wts[s];
        weights[2 * i + 1] = wts[s];
    }
}

// Kii: 2^d * GL16^d of the IE integrand on [0,h/2]^d
// (1D/codes/integral1d.hpp:17-39, 2D/codes/integral2d.hpp:60-90, 3D/codes/integral3d.hpp:13-50)
double ie_kii(int dim, double h) {
    double nd[16], w[16];
    gauss_legendre16(nd, w);
    double a = 0.0, b = 0.5 * h;
    double cx = 0.5 * (b - a), Lx = 0.5 * (b + a);
    double res = 0.0;
    if (dim == 1) {
        for (int i = 0; i < 16; ++i) res += w[i] * std::log(std::fabs(1.0 * (cx * nd[i] + Lx)));
        res *= cx;
        res *= 2;
    } else if (dim == 2) {
        double c = -1.0 / (2 * PI);
        for (int i = 0; i < 16; ++i) {
            double tx = cx * nd[i] + Lx;
            for (int j = 0; j < 16; ++j) {
                double ty = cx * nd[j] + Lx;
                res += w[i] * w[j] * (c * 0.5 * std::log(tx * tx + ty * ty));
            }
        }
        res *= (cx * cx);
        res *= 4;
    } else {
        for (int i = 0; i < 16; ++i) {
            double tx = cx * nd[i] + Lx;
            for (int j = 0; j < 16; ++j) {
                double ty = cx * nd[j] + Lx;
                for (int k = 0; k < 16; ++k) {
                    double tz = cx * nd[k] + Lx;
                    res += (w[i] * w[j] * w[k] * (1.0 / ((4 * PI) * std::sqrt(tx * tx + ty * ty + tz * tz))));
                }
            }
        }
        res *= (cx * cx * cx);
        res *= 8;
    }
    return res;
}

// ---------------------------------------------------------------- kernel
struct Kernel {
    int dim = 3, choice = 0;
    int N = 0;
    std::vector<double> x, y, z;  // input order
    double diag = 0.0, Kii = 0.0, h = 0.0;

    void init(int compat_1d_dup) {
        if (choice == 0) {
            if (dim == 1) {
                h = 1.0 / double(N);
                Kii = ie_kii(1, h);
                diag = (1.0 + Kii) / h;  // 1D/codes/kernel.hpp:108-114
            } else if (dim == 2) {
                h = 1.0 / std::sqrt(double(N));
                Kii = ie_kii(2, h);
                diag = (1.0 + Kii) / std::pow(h, 2.0);  // 2D/codes/kernel.hpp:101-122
            } else {
                h = 1.0 / std::cbrt(double(N));
                Kii = ie_kii(3, h);
                diag = (1.0 + Kii) / std::pow(h, 3.0);  // 3D/codes/kernel.hpp:101-129
            }
        } else {
            // 1D: the particle vector holds 2N entries (1D/codes/matvec.hpp:35,38-39)
            if (dim == 1) diag = std::sqrt(double(compat_1d_dup ? 2 * N : N));
            else if (dim == 2) diag = 1.0;          // 2D/codes/kernel.hpp:312-313
            else diag = std::sqrt(double(N));        // 3D/codes/kernel.hpp:186-187
        }
    }
    double operator()(int i, int j) const {
        if (i == j) return diag;
        const double a = 0.0001;
        if (dim == 1) {
            double R = std::fabs(x[i] - x[j]);
            if (choice == 0) return std::log(R);
            return R < a ? R / a : a / R;
        } else if (dim == 2) {
            double dx = x[i] - x[j], dy = y[i] - y[j];
            double R2 = dx * dx + dy * dy;
            if (choice == 0) return (-1.0 / (2 * PI)) * 0.5 * std::log(R2);
            double R = std::sqrt(R2);
            return R < a ? R / a : a / R;
        } else {
            double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            double R2 = dx * dx + dy * dy + dz * dz;
            double R = std::sqrt(R2);
            if (choice == 0) return 1.0 / (4.0 * PI * R);
            if (R < a) return (R * std::log(R) - 1.0) / (a * std::log(a) - 1.0);
            return (1.0 * std::log(R)) / std::log(a);
        }
    }
    void row(int i, const std::vector<int> &cols, Vec &out) const {
        int n = (int)cols.size();
        out.resize(n);
#pragma omp parallel for
        for (int k = 0; k < n; ++k) out[k] = (*this)(i, cols[k]);
    }
    void col(int j, const std::vector<int> &rows, Vec &out) const {
        int n = (int)rows.size();
        out.resize(n);
#pragma omp parallel for
        for (int k = 0; k < n; ++k) out[k] = (*this)(rows[k], j);
    }
    Mat block(const std::vector<int> &rows, const std::vector<int> &cols) const {
        Mat M((int)rows.size(), (int)cols.size());
        for (int j = 0; j < M.r; ++j)
            for (int k = 0; k < M.c; ++k) M(j, k) = (*this)(rows[j], cols[k]);
        return M;
    }
};

// ---------------------------------------------------------------- progress
// With H2W_ORACLE_PROGRESS set, long runs report what they finished (levels, large crosses)
// on stderr as they go.
bool progress_on() {
    static const bool on = std::getenv("H2W_ORACLE_PROGRESS") != nullptr;
    return on;
}
double wall() { return omp_get_wtime(); }

// ---------------------------------------------------------------- ACA
// First index, in ascending order over the remaining set, of max |v|.
// (2D/codes/aca.hpp:23-38); returns -1 on an empty set.
int max_abs(const Vec &v, const std::vector<char> &rem) {
    int idx = -1;
    double mx = 0.0;
    for (size_t i = 0; i < v.size(); ++i) {
        if (!rem[i]) continue;
        if (idx < 0) {
            idx = (int)i;
            mx = std::fabs(v[i]);
        } else if (std::fabs(v[i]) > mx) {
            idx = (int)i;
            mx = std::fabs(v[i]);
        }
    }
    return idx;
}

struct AcaResult {
    int rank = 0;
    std::vector<int> rb, cb;  // local pivot indices
    std::vector<Vec> U, V, Acv, Arv;
    Mat L, R;
};

// Partial-pivot ACA (2D/codes/aca.hpp:296-446 with Ac/Ar, :158-294 CUR,
// 1D/codes/aca.hpp:40-156 UV): the three variants run the same pivot loop.
void aca(const Kernel &K, const std::vector<int> &rows, const std::vector<int> &cols, int tol_pow,
         AcaResult &o, bool keep_acar) {
    const double tol = std::pow(10, -1.0 * tol_pow);
    int N1 = (int)rows.size(), N2 = (int)cols.size();
    o = AcaResult();
    if (N1 == 0 || N2 == 0) return;  // the reference's behaviour here is undefined
    int mn = std::min(N1, N2);
    std::vector<char> rrem(N1, 1), crem(N2, 1);
    Vec row, col, u(N1), v, row_temp, col_temp;
    int l, ci = -1;
    for (l = 0; l < N1; ++l) {
        K.row(rows[l], cols, row);
        ci = max_abs(row, crem);
        if (std::fabs(row[ci]) > std::pow(10, -36)) break;
    }
    if (l == N1) {
        o.L = Mat(0, 0);
        o.R = Mat(0, 0);
        return;
    }
    int ri = l;
    v = row;
    o.rb.push_back(ri);
    o.cb.push_back(ci);
    K.col(cols[ci], rows, col);
    for (int t = 0; t < N1; ++t) u[t] = col[t] / row[ci];
    o.U.push_back(u);
    o.V.push_back(v);
    if (keep_acar) {
        o.Acv.push_back(col);
        o.Arv.push_back(row);
    }
    crem[ci] = 0;
    rrem[ri] = 0;
    int rank = 1;
    double normS = 0.0;
    ri = max_abs(col, rrem);
    while (vnorm(u) * vnorm(v) > tol * normS && rank < mn) {
        o.rb.push_back(ri);
        K.row(rows[ri], cols, row);
        row_temp = row;
        for (int q = 0; q < rank; ++q) {
            double uq = o.U[q][ri];
            const Vec &Vq = o.V[q];
            for (int s = 0; s < N2; ++s) row[s] -= uq * Vq[s];
        }
        ci = max_abs(row, crem);
        o.cb.push_back(ci);
        v = row;
        K.col(cols[ci], rows, col);
        col_temp = col;
        for (int q = 0; q < rank; ++q) {
            double vq = o.V[q][ci];
            const Vec &Uq = o.U[q];
            for (int s = 0; s < N1; ++s) col[s] -= vq * Uq[s];
        }
        for (int t = 0; t < N1; ++t) u[t] = col[t] / row[ci];
        double un = vnorm(u), vn = vnorm(v);
        if (un < 1e-36 || vn < 1e-36) {
            o.rb.pop_back();
            o.cb.pop_back();
            break;
        }
        o.U.push_back(u);
        o.V.push_back(v);
        if (keep_acar) {
            o.Acv.push_back(col_temp);
            o.Arv.push_back(row_temp);
        }
        ++rank;
        crem[ci] = 0;
        rrem[ri] = 0;
        normS = std::pow(normS, 2) + std::pow(un * vn, 2);
        for (int q = 0; q < rank; ++q) normS += 2 * std::fabs(vdot(o.U[q], u) * vdot(o.V[q], v));
        normS = std::sqrt(normS);
        ri = max_abs(col, rrem);
        if (progress_on() && (double)N1 * N2 > 1e8 && rank % 100 == 0) {
            std::fprintf(stderr, "[oracle]   cross %d x %d: rank %d\n", N1, N2, rank);
            std::fflush(stderr);
        }
    }
    o.rank = rank;
    o.L = Mat(rank, rank);
    o.R = Mat(rank, rank);
    for (int i = 0; i < rank; ++i) {
        o.L(i, i) = 1.0;
        for (int j = 0; j < i; ++j) o.L(i, j) = o.U[j][o.rb[i]];
    }
    for (int i = 0; i < rank; ++i) {
        o.R(i, i) = o.V[i][o.cb[i]];
        for (int j = 0; j < i; ++j) o.R(j, i) = o.V[j][o.cb[i]];
    }
}

// ---------------------------------------------------------------- tree
struct Box {
    double c[3] = {0, 0, 0};
    int ic[3] = {0, 0, 0};             // integer coordinates at its level
    int nbr[26];                        // near neighbour slots (-1 = none)
    std::vector<int> il;                // interaction list, reference order
    std::vector<char> vert;             // isVertex per il entry
    std::vector<int> pts;               // chargeLocations (input-order indices)
    // far (nested, both variants)
    std::vector<int> Ir, Ic;
    Mat L2P;
    std::map<int, Mat> M2L_far;
    // vertex nested
    std::vector<int> Irv, Icv;
    Mat Lv, Rv;
    std::vector<Mat> M2Mv;
    std::map<int, Mat> M2L_ver;
    // semi-nested CUR per il index
    std::map<int, Mat> Lc, Rc, Ac, Ar;
    std::map<int, std::vector<int>> rbas, cbas;
    // 1D semi UV
    std::map<int, Mat> U, Vt;
    // near
    std::vector<Mat> dense;  // one per nbr slot + self (last)
    // matvec state
    Vec charges, outgoing, incoming, incoming_ver, potential, potential_ver;
    bool has_far_rank = false;
    Box() {
        for (int i = 0; i < 26; ++i) nbr[i] = -1;
    }
};

struct Oracle {
    int dim = 3, variant = 0, nL = 1, n_max = 64, tol_pow = 10;
    double L = 1.0;
    int nchild = 8, nslot = 26;
    Kernel K;
    std::vector<std::vector<Box>> tree;
    std::vector<int> nBoxes;
    std::vector<double> radius;
    double memory = 0.0;
    bool assembled = false;

    // child c -> bit offsets (x,y,z); numbering as 3D/codes/HODLR3Dtree.hpp:1700-1741,
    // 2D/codes/HODLR2Dtree.hpp:426-449
    void child_bits(int c, int *b) const {
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
    // neighbour slot of an integer offset; -1 for a vertex offset / self.
    // 3D: n = 9(dz+1)+3(dy+1)+(dx+1), minus one above 13 (3D/codes/HODLR3Dtree.hpp:163-169);
    // 2D: 0 below, 1 right, 2 above, 3 left (2D/codes/HODLR2Dtree.hpp:142-205)
    int slot_of(const int *d) const {
        if (dim == 2) {
            if (d[0] == 0 && d[1] == -1) return 0;
            if (d[0] == 1 && d[1] == 0) return 1;
            if (d[0] == 0 && d[1] == 1) return 2;
            if (d[0] == -1 && d[1] == 0) return 3;
            return -1;
        }
        int nz = (d[0] != 0) + (d[1] != 0) + (d[2] != 0);
        if (nz == 0 || nz == 3) return -1;
        int n = 9 * (d[2] + 1) + 3 * (d[1] + 1) + (d[0] + 1);
        return n > 13 ? n - 1 : n;
    }

    void build(const double *pts, int64_t N, int nlev, int compat) {
        nchild = 1 << dim;
        nslot = dim == 3 ? 26 : (dim == 2 ? 4 : 2);
        K.dim = dim;
        K.N = (int)N;
        K.x.assign(N, 0.0);
        K.y.assign(N, 0.0);
        K.z.assign(N, 0.0);
        if (pts) {
            for (int64_t i = 0; i < N; ++i) {
                K.x[i] = pts[i];
                if (dim > 1) K.y[i] = pts[N + i];
                if (dim > 2) K.z[i] = pts[2 * N + i];
            }
        } else {
            // reference grids: 1D/codes/HODLR1Dtree.hpp:59-65, 2D/codes/HODLR2Dtree.hpp:75-91,
            // 3D/codes/HODLR3Dtree.hpp:82-102
            if (dim == 1) {
                for (int k = 0; k < N; ++k) K.x[k] = -1.0 + 2.0 * (k + 1.0) / (N + 1.0);
            } else if (dim == 2) {
                int n = (int)std::llround(std::sqrt((double)N));
                std::vector<double> nd(n);
                for (int k = 0; k < n; ++k) nd[k] = -L + 2.0 * L * (k + 1.0) / (n + 1.0);
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k) {
                        K.x[(size_t)j * n + k] = nd[k];
                        K.y[(size_t)j * n + k] = nd[j];
                    }
            } else {
                int n = (int)std::llround(std::cbrt((double)N));
                std::vector<double> nd(n);
                for (int k = 0; k < n; ++k) nd[k] = -L + 2.0 * L * (k + 1.0) / (n + 1.0);
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k)
                        for (int i = 0; i < n; ++i) {
                            size_t p = ((size_t)j * n + k) * n + i;
                            K.x[p] = nd[k];
                            K.y[p] = nd[j];
                            K.z[p] = nd[i];
                        }
            }
        }
        K.init(compat);
        nL = nlev;
        nBoxes.assign(nL + 1, 1);
        radius.assign(nL + 1, L);
        for (int k = 1; k <= nL; ++k) {
            nBoxes[k] = nBoxes[k - 1] * nchild;
            radius[k] = 0.5 * radius[k - 1];
        }
        tree.assign(nL + 1, std::vector<Box>());
        for (int j = 0; j <= nL; ++j) tree[j].resize(nBoxes[j]);
        // centres and integer coordinates
        for (int j = 0; j < nL; ++j) {
            double shift = 0.5 * radius[j];
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &P = tree[j][k];
                for (int c = 0; c < nchild; ++c) {
                    Box &C = tree[j + 1][nchild * k + c];
                    int b[3];
                    child_bits(c, b);
                    for (int d = 0; d < dim; ++d) {
                        C.c[d] = b[d] ? P.c[d] + shift : P.c[d] - shift;
                        C.ic[d] = 2 * P.ic[d] + b[d];
                    }
                }
            }
        }
        if (dim == 1) build_lists_1d();
        else build_lists();
        // points to boxes (3D/codes/HODLR3Dtree.hpp:1769-1852)
        tree[0][0].pts.resize(N);
        for (int i = 0; i < N; ++i) tree[0][0].pts[i] = i;
        for (int j = 0; j < nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &P = tree[j][k];
                for (int idx : P.pts) {
                    int c;
                    if (dim == 1) c = K.x[idx] <= P.c[0] ? 0 : 1;
                    else if (dim == 2) {
                        bool xl = K.x[idx] <= P.c[0], yl = K.y[idx] <= P.c[1];
                        c = xl ? (yl ? 0 : 3) : (yl ? 1 : 2);
                    } else {
                        bool zl = K.z[idx] <= P.c[2], xl = K.x[idx] <= P.c[0], yl = K.y[idx] <= P.c[1];
                        c = xl ? (yl ? 0 : 3) : (yl ? 1 : 2);
                        if (!zl) c += 4;
                    }
                    tree[j + 1][nchild * k + c].pts.push_back(idx);
                }
            }
        for (int j = nL - 1; j >= 1; --j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &P = tree[j][k];
                P.pts.clear();
                for (int c = 0; c < nchild; ++c) {
                    Box &C = tree[j + 1][nchild * k + c];
                    P.pts.insert(P.pts.end(), C.pts.begin(), C.pts.end());
                }
            }
    }

    // 1D: the only partner is the sibling, the near field is the box itself
    // (1D/codes/HODLR1Dtree.hpp:116-147,466-484)
    void build_lists_1d() {
        for (int j = 1; j <= nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                int sib = (k % 2 == 0) ? k + 1 : k - 1;
                B.nbr[k % 2 == 0 ? 1 : 0] = sib;
                B.il.push_back(sib);
                B.vert.push_back(1);
            }
    }

    // 2D/3D: restates the hand tables by their rule: near siblings and
    // near children of the parent's neighbours go to neighbour slots, the rest
    // to the interaction list, parent neighbours taken in slot order.
    void build_lists() {
        for (int j = 0; j < nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &P = tree[j][k];
                for (int c = 0; c < nchild; ++c) {
                    int nC = nchild * k + c;
                    Box &B = tree[j + 1][nC];
                    // siblings
                    for (int c2 = 0; c2 < nchild; ++c2) {
                        if (c2 == c) continue;
                        Box &S = tree[j + 1][nchild * k + c2];
                        int d[3] = {0, 0, 0};
                        for (int q = 0; q < dim; ++q) d[q] = S.ic[q] - B.ic[q];
                        int s = slot_of(d);
                        if (s >= 0) B.nbr[s] = nchild * k + c2;
                        else B.il.push_back(nchild * k + c2);
                    }
                    for (int s = 0; s < nslot; ++s) {
                        int nN = P.nbr[s];
                        if (nN == -1) continue;
                        std::vector<int> order;
                        if (dim == 3) {
                            // vertex child first, then the rest ascending (3D/codes/HODLR3Dtree.hpp:172-302)
                            int vc = -1;
                            for (int c2 = 0; c2 < nchild; ++c2) {
                                Box &C = tree[j + 1][nchild * nN + c2];
                                int d[3];
                                for (int q = 0; q < 3; ++q) d[q] = C.ic[q] - B.ic[q];
                                if (std::abs(d[0]) == 1 && std::abs(d[1]) == 1 && std::abs(d[2]) == 1) vc = c2;
                            }
                            if (vc >= 0) order.push_back(vc);
                            for (int c2 = 0; c2 < nchild; ++c2)
                                if (c2 != vc) order.push_back(c2);
                        } else {
                            for (int c2 = 0; c2 < nchild; ++c2) order.push_back(c2);
                        }
                        for (int c2 : order) {
                            int id = nchild * nN + c2;
                            Box &C = tree[j + 1][id];
                            int d[3] = {0, 0, 0};
                            int mx = 0;
                            for (int q = 0; q < dim; ++q) {
                                d[q] = C.ic[q] - B.ic[q];
                                mx = std::max(mx, std::abs(d[q]));
                            }
                            int sl = mx <= 1 ? slot_of(d) : -1;
                            if (sl >= 0) B.nbr[sl] = id;
                            else B.il.push_back(id);
                        }
                    }
                }
            }
        // vertex flags by centre distance (2D/codes/HODLR2Dtree.hpp:452-475,
        // 3D/codes/HODLR3Dtree.hpp:1744-1767)
        for (int j = 1; j <= nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                B.vert.assign(B.il.size(), 0);
                for (size_t i = 0; i < B.il.size(); ++i) {
                    Box &F = tree[j][B.il[i]];
                    double s = 0.0;
                    for (int q = 0; q < dim; ++q) s += (B.c[q] - F.c[q]) * (B.c[q] - F.c[q]);
                    double t = std::sqrt(s);
                    B.vert[i] = (t - (2.0 * std::sqrt((double)dim) * radius[j]) <= 1.0e-16) ? 1 : 0;
                }
            }
    }

    // --------------------------------------------------------- assembly
    void row_candidates_far(int j, int k, std::vector<int> &out) {
        if (j == nL) {
            out.insert(out.end(), tree[j][k].pts.begin(), tree[j][k].pts.end());
        } else {
            for (int c = 0; c < nchild; ++c) {
                Box &C = tree[j + 1][nchild * k + c];
                out.insert(out.end(), C.Ir.begin(), C.Ir.end());
            }
        }
    }
    // M2L blocks K(Ir_B, Ir_F) of one level, stored for both directions (3D/codes/HODLR3Dtree.hpp:1978-1998,
    // 2164-2185).  Each unordered pair is evaluated once, by its smaller box, and the transpose is
    // copied to the other box.  Boxes run in parallel (assembly is not what the CPU baseline times);
    // every box map is written by one thread only.
    void assemble_m2l(int j, bool ver) {
        const int nb = nBoxes[j];
        std::vector<std::vector<std::pair<int, Mat>>> up(nb);
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < nb; ++k) {
            Box &B = tree[j][k];
            for (size_t i = 0; i < B.il.size(); ++i) {
                int f = B.il[i];
                if ((bool)B.vert[i] != ver || f < k) continue;
                up[k].push_back({f, K.block(ver ? B.Irv : B.Ir, ver ? tree[j][f].Irv : tree[j][f].Ir)});
            }
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < nb; ++k) {
            Box &B = tree[j][k];
            auto &dst = ver ? B.M2L_ver : B.M2L_far;
            for (auto &e : up[k]) dst[e.first] = e.second;
            for (size_t i = 0; i < B.il.size(); ++i) {
                int f = B.il[i];
                if ((bool)B.vert[i] != ver || f >= k) continue;
                bool found = false;
                for (auto &e : up[f])
                    if (e.first == k) {
                        dst[f] = transpose(e.second);
                        found = true;
                    }
                if (!found) dst[f] = K.block(ver ? B.Irv : B.Ir, ver ? tree[j][f].Irv : tree[j][f].Ir);
            }
        }
    }
    void assemble_far() {
        if (dim == 1) return;
        // boxes of one level are independent (the reference loops over them serially; the pivots do
        // not depend on the order, the row/column gathers inside the cross then run serially)
        for (int j = nL; j >= 2; --j) {
#pragma omp parallel for schedule(dynamic, 1)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                std::vector<int> rows, cols;
                row_candidates_far(j, k, rows);
                for (size_t i = 0; i < B.il.size(); ++i)
                    if (!B.vert[i]) row_candidates_far(j, B.il[i], cols);
                AcaResult a;
                aca(K, rows, cols, tol_pow, a, true);
                if (a.rank > 0) {
                    for (int r : a.rb) B.Ir.push_back(rows[r]);
                    for (int c : a.cb) B.Ic.push_back(cols[c]);
                    Mat Ac((int)rows.size(), a.rank);
                    for (int q = 0; q < a.rank; ++q)
                        for (int t = 0; t < Ac.r; ++t) Ac(t, q) = a.Acv[q][t];
                    B.L2P = solve_right_lower(a.L, solve_right_upper(a.R, Ac));
                } else {
                    B.L2P = Mat((int)rows.size(), 0);
                }
            }
            note("far cross", j);
        }
        for (int j = 2; j <= nL; ++j) {
            assemble_m2l(j, false);
            note("far M2L", j);
        }
    }
    void assemble_vertex_nested() {
        for (int j = 1; j <= nL; ++j) {
#pragma omp parallel for schedule(dynamic, 1)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                std::vector<int> cols;
                for (size_t i = 0; i < B.il.size(); ++i)
                    if (B.vert[i]) {
                        Box &V = tree[j][B.il[i]];
                        cols.insert(cols.end(), V.pts.begin(), V.pts.end());
                    }
                if (j > 1) {
                    Box &P = tree[j - 1][k / nchild];
                    cols.insert(cols.end(), P.Icv.begin(), P.Icv.end());
                }
                AcaResult a;
                aca(K, B.pts, cols, tol_pow, a, false);
                B.Lv = a.L;
                B.Rv = a.R;
                for (int r : a.rb) B.Irv.push_back(B.pts[r]);
                for (int c : a.cb) B.Icv.push_back(cols[c]);
            }
            note("vertex cross", j);
        }
        for (int j = 1; j <= nL; ++j) {
            assemble_m2l(j, true);
            note("vertex M2L", j);
        }
        for (int j = 1; j <= nL; ++j)
#pragma omp parallel for schedule(dynamic, 1)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                if (j == nL) {
                    B.M2Mv.assign(1, solve_lr_transpose(B.Lv, B.Rv, K.block(B.Icv, B.pts)));
                } else {
                    B.M2Mv.resize(nchild);
                    for (int c = 0; c < nchild; ++c)
                        B.M2Mv[c] = solve_lr_transpose(B.Lv, B.Rv, K.block(B.Icv, tree[j + 1][nchild * k + c].Irv));
                }
            }
    }
    void assemble_vertex_semi() {
        for (int j = 1; j <= nL; ++j)
#pragma omp parallel for schedule(dynamic, 1)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                for (size_t i = 0; i < B.il.size(); ++i) {
                    if (!B.vert[i] && variant != 2) continue;
                    int ki = B.il[i];
                    Box &V = tree[j][ki];
                    AcaResult a;
                    aca(K, B.pts, V.pts, tol_pow, a, false);
                    if (dim == 1) {
                        // ACA_only_nodesUV: U (m x r), Vt (r x n)
                        Mat U((int)B.pts.size(), a.rank), Vt(a.rank, (int)V.pts.size());
                        for (int q = 0; q < a.rank; ++q) {
                            for (int t = 0; t < U.r; ++t) U(t, q) = a.U[q][t];
                            for (int t = 0; t < Vt.c; ++t) Vt(q, t) = a.V[q][t];
                        }
                        B.U[ki] = U;
                        B.Vt[ki] = Vt;
                    } else {
                        B.Lc[(int)i] = a.L;
                        B.Rc[(int)i] = a.R;
                        std::vector<int> rb, cbv;
                        for (int q = 0; q < a.rank; ++q) {
                            rb.push_back(B.pts[a.rb[q]]);
                            cbv.push_back(V.pts[a.cb[q]]);
                        }
                        B.rbas[(int)i] = rb;
                        B.cbas[(int)i] = cbv;
                        if (a.rank > 0) {
                            B.Ac[ki] = K.block(B.pts, cbv);
                            B.Ar[ki] = K.block(rb, V.pts);
                        }
                    }
                }
            }
    }
    void assemble_near() {
        int nb = nBoxes[nL];
#pragma omp parallel for
        for (int k = 0; k < nb; ++k) {
            Box &B = tree[nL][k];
            B.dense.assign(nslot + 1, Mat());
            if (dim > 1)
                for (int n = 0; n < nslot; ++n)
                    if (B.nbr[n] != -1) B.dense[n] = K.block(B.pts, tree[nL][B.nbr[n]].pts);
            B.dense[nslot] = K.block(B.pts, B.pts);
        }
    }
    void note(const char *what, int j) const {
        if (progress_on()) {
            std::fprintf(stderr, "[oracle] %s level %d done (%.1f s)\n", what, j, wall() - t_start);
            std::fflush(stderr);
        }
    }
    double t_start = 0.0;
    void assemble() {
        t_start = wall();
        if (variant == 0) {
            assemble_far();
            assemble_vertex_nested();
        } else if (variant == 1) {
            assemble_far();
            assemble_vertex_semi();
        } else {
            // non-nested H* (USE_HODLRdD): CUR on every interaction-list pair, no far NCA
            // (2D/codes/HODLR2Dtree.hpp:971-980,1007-1021; 3D/codes/HODLR3Dtree.hpp:2239-2248,2282-2296)
            assemble_vertex_semi();
        }
        assemble_near();
        note("near field", nL);
        find_memory();
        assembled = true;
    }

    // --------------------------------------------------------- storage
    void find_memory() {
        memory = 0.0;
        if (dim == 1) {
            // 1D/codes/HODLR1Dtree.hpp:535-588
            if (variant == 0) {
                for (int j = 1; j <= nL; ++j)
                    for (int k = 0; k < nBoxes[j]; ++k)
                        for (auto &M : tree[j][k].M2Mv) memory += 2 * (double)M.size();
                for (int j = 1; j <= nL; ++j)
                    for (int k = 0; k < nBoxes[j]; ++k)
                        for (auto &kv : tree[j][k].M2L_ver) memory += (double)kv.second.size();
                for (int k = 0; k < nBoxes[nL]; ++k) memory += (double)tree[nL][k].dense[nslot].size();
            } else {
                for (int j = 1; j <= nL; ++j)
                    for (int k = 0; k < nBoxes[j]; ++k)
                        for (auto &kv : tree[j][k].U)
                            memory += (double)kv.second.size() + (double)tree[j][k].Vt[kv.first].size();
            }
            return;
        }
        if (variant == 2) {
            // findMemory_nn (3D/codes/HODLR3Dtree.hpp:2561-2583): L + R + Ac + Ar of every pair, near field
            for (int j = 1; j <= nL; ++j)
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    for (size_t i = 0; i < B.il.size(); ++i) {
                        int ki = B.il[i];
                        memory += (double)B.Lc[(int)i].size() + (double)B.Rc[(int)i].size();
                        if (B.Ac.count(ki)) memory += (double)B.Ac[ki].size() + (double)B.Ar[ki].size();
                    }
                }
            for (int k = 0; k < nBoxes[nL]; ++k) {
                Box &B = tree[nL][k];
                memory += (double)B.pts.size() * B.pts.size();
                for (int n = 0; n < nslot; ++n)
                    if (B.nbr[n] != -1) memory += (double)B.pts.size() * tree[nL][B.nbr[n]].pts.size();
            }
            return;
        }
        for (int j = 2; j <= nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) memory += 2 * (double)tree[j][k].L2P.size();
        if (variant == 0)
            for (int j = 1; j <= nL; ++j)
                for (int k = 0; k < nBoxes[j]; ++k)
                    for (auto &M : tree[j][k].M2Mv) memory += 2 * (double)M.size();
        for (int j = 2; j <= nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                for (size_t i = 0; i < B.il.size(); ++i)
                    if (!B.vert[i]) memory += (double)B.M2L_far[B.il[i]].size();
            }
        for (int j = 1; j <= nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                for (size_t i = 0; i < B.il.size(); ++i) {
                    if (!B.vert[i]) continue;
                    int ki = B.il[i];
                    if (variant == 0) memory += (double)B.M2L_ver[ki].size();
                    else {
                        memory += (double)B.Lc[(int)i].size() + (double)B.Rc[(int)i].size();
                        if (B.Ac.count(ki)) memory += (double)B.Ac[ki].size() + (double)B.Ar[ki].size();
                    }
                }
            }
        for (int k = 0; k < nBoxes[nL]; ++k) {
            Box &B = tree[nL][k];
            memory += (double)B.pts.size() * B.pts.size();
            for (int n = 0; n < nslot; ++n)
                if (B.nbr[n] != -1) memory += (double)B.pts.size() * tree[nL][B.nbr[n]].pts.size();
        }
    }

    // --------------------------------------------------------- matvec
    void matvec(const double *q, double *out) {
        int N = K.N;
        for (int j = 1; j <= nL; ++j)
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                B.charges.resize(B.pts.size());
                for (size_t i = 0; i < B.pts.size(); ++i) B.charges[i] = q[B.pts[i]];
                B.potential.clear();
                B.potential_ver.clear();
            }
        // far up / M2L / down (3D/codes/HODLR3Dtree.hpp:1941-2055)
        if (dim > 1 && variant != 2) {
            for (int j = nL; j >= 2; --j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    Vec src;
                    if (j == nL) src = B.charges;
                    else
                        for (int c = 0; c < nchild; ++c) {
                            Box &C = tree[j + 1][nchild * k + c];
                            src.insert(src.end(), C.outgoing.begin(), C.outgoing.end());
                        }
                    B.outgoing.assign(B.L2P.c, 0.0);
                    gemv_t(B.L2P, src.data(), B.outgoing.data(), false);
                }
            }
            for (int j = 2; j <= nL; ++j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    B.incoming.assign(B.Ir.size(), 0.0);
                    for (size_t i = 0; i < B.il.size(); ++i) {
                        if (B.vert[i]) continue;
                        int f = B.il[i];
                        gemv(B.M2L_far[f], tree[j][f].outgoing.data(), B.incoming.data(), true);
                    }
                }
            }
            for (int j = 2; j <= nL; ++j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    Vec t(B.L2P.r, 0.0);
                    gemv(B.L2P, B.incoming.data(), t.data(), false);
                    if (j != nL) {
                        int ind = 0;
                        for (int c = 0; c < nchild; ++c) {
                            Box &C = tree[j + 1][nchild * k + c];
                            for (size_t s = 0; s < C.Ir.size(); ++s) C.incoming[s] += t[ind + s];
                            ind += (int)C.Ir.size();
                        }
                    } else {
                        B.potential = t;
                    }
                }
            }
        }
        if (variant == 0) {
            // vertex up / M2L / down (3D/codes/HODLR3Dtree.hpp:2135-2237)
            for (int j = nL; j >= 1; --j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    B.outgoing.assign(B.Irv.size(), 0.0);
                    if (j == nL) gemv(B.M2Mv[0], B.charges.data(), B.outgoing.data(), false);
                    else
                        for (int c = 0; c < nchild; ++c)
                            gemv(B.M2Mv[c], tree[j + 1][nchild * k + c].outgoing.data(), B.outgoing.data(), true);
                }
            }
            for (int j = 1; j <= nL; ++j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    B.incoming_ver.assign(B.Irv.size(), 0.0);
                    for (size_t i = 0; i < B.il.size(); ++i) {
                        if (!B.vert[i]) continue;
                        int f = B.il[i];
                        gemv(B.M2L_ver[f], tree[j][f].outgoing.data(), B.incoming_ver.data(), true);
                    }
                }
            }
            for (int j = 1; j <= nL; ++j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    if (j != nL) {
                        for (int c = 0; c < nchild; ++c) {
                            Box &C = tree[j + 1][nchild * k + c];
                            gemv_t(B.M2Mv[c], B.incoming_ver.data(), C.incoming_ver.data(), true);
                        }
                    } else {
                        B.potential_ver.assign(B.pts.size(), 0.0);
                        gemv_t(B.M2Mv[0], B.incoming_ver.data(), B.potential_ver.data(), false);
                    }
                }
            }
        } else {
            // semi-nested vertex blocks at every level (3D/codes/HODLR3Dtree.hpp:2323-2351,
            // 1D/codes/HODLR1Dtree.hpp:445-463)
            for (int j = 1; j <= nL; ++j) {
#pragma omp parallel for
                for (int k = 0; k < nBoxes[j]; ++k) {
                    Box &B = tree[j][k];
                    B.potential_ver.assign(B.pts.size(), 0.0);
                    for (size_t i = 0; i < B.il.size(); ++i) {
                        if (!B.vert[i] && variant != 2) continue;
                        int ki = B.il[i];
                        Box &V = tree[j][ki];
                        if (dim == 1) {
                            Mat &U = B.U[ki], &Vt = B.Vt[ki];
                            Vec t(Vt.r, 0.0);
                            gemv(Vt, V.charges.data(), t.data(), false);
                            gemv(U, t.data(), B.potential_ver.data(), true);
                        } else {
                            if (B.cbas[(int)i].empty()) continue;
                            Vec t(B.Ar[ki].r, 0.0);
                            gemv(B.Ar[ki], V.charges.data(), t.data(), false);
                            solve_lr_vec(B.Lc[(int)i], B.Rc[(int)i], t.data());
                            gemv(B.Ac[ki], t.data(), B.potential_ver.data(), true);
                        }
                    }
                }
            }
        }
        // near field (3D/codes/HODLR3Dtree.hpp:2392-2423)
        {
            int nb = nBoxes[nL];
#pragma omp parallel for
            for (int k = 0; k < nb; ++k) {
                Box &B = tree[nL][k];
                if (B.potential.size() != B.pts.size()) B.potential.assign(B.pts.size(), 0.0);
                if (dim > 1)
                    for (int n = 0; n < nslot; ++n)
                        if (B.nbr[n] != -1)
                            gemv(B.dense[n], tree[nL][B.nbr[n]].charges.data(), B.potential.data(), true);
                gemv(B.dense[nslot], B.charges.data(), B.potential.data(), true);
            }
        }
        // collect in tree order, then reorder (3D/codes/HODLR3Dtree.hpp:2426-2503)
        Vec pot(N, 0.0);
        for (int j = 1; j <= nL; ++j) {
            Vec tmp(N, 0.0);
            int count = 0;
            for (int k = 0; k < nBoxes[j]; ++k) {
                Box &B = tree[j][k];
                size_t n = B.pts.size();
                for (size_t i = 0; i < n; ++i) {
                    double v = 0.0;
                    if (j == nL) v += B.potential[i];
                    if (B.potential_ver.size() == n) v += B.potential_ver[i];
                    tmp[count + i] = v;
                }
                count += (int)n;
            }
            for (int i = 0; i < N; ++i) pot[i] += tmp[i];
        }
        int count = 0;
        for (int k = 0; k < nBoxes[nL]; ++k)
            for (int idx : tree[nL][k].pts) out[idx] = pot[count++];
    }

    void dense_matvec(const double *q, double *out, const int64_t *rows, int64_t nrows) {
        int N = K.N;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t t = 0; t < nrows; ++t) {
            int i = rows ? (int)rows[t] : (int)t;
            double s = 0.0;
            for (int j = 0; j < N; ++j) s += K(i, j) * q[j];
            out[t] = s;
        }
    }

    // --------------------------------------------------------- GMRES
    // Unrestarted GMRES with MGS and Givens rotations (1D/codes/gmres.hpp:30-200)
    static void gen_rot(double dx, double dy, double &cs, double &sn) {
        if (dy == 0.0) {
            cs = 1.0;
            sn = 0.0;
        } else if (std::fabs(dy) > std::fabs(dx)) {
            double t = dx / dy;
            sn = 1.0 / std::sqrt(1.0 + t * t);
            cs = t * sn;
        } else {
            double t = dy / dx;
            cs = 1.0 / std::sqrt(1.0 + t * t);
            sn = t * cs;
        }
    }
    static void app_rot(double &dx, double &dy, double cs, double sn) {
        double t = cs * dx + sn * dy;
        dy = -sn * dx + cs * dy;
        dx = t;
    }
    int gmres(const double *b, double *x, int max_iter, double tol, double *resid_out, int *matvecs) {
        int N = K.N, m = max_iter;
        std::vector<double> H((size_t)(m + 1) * m, 0.0);
        auto h = [&](int i, int j) -> double & { return H[(size_t)j * (m + 1) + i]; };
        Vec s(m + 1, 0.0), cs(m + 1, 0.0), sn(m + 1, 0.0), w(N), r(N);
        double normb = 0.0;
        for (int i = 0; i < N; ++i) normb += b[i] * b[i];
        normb = std::sqrt(normb);
        int nmv = 0;
        matvec(x, w.data());
        ++nmv;
        for (int i = 0; i < N; ++i) r[i] = b[i] - w[i];
        double beta = vnorm(r);
        if (normb == 0.0) normb = 1;
        double resid = beta / normb;
        if (resid <= tol) {
            *resid_out = resid;
            if (matvecs) *matvecs = nmv;
            return 0;
        }
        std::vector<Vec> v(m + 1);
        int j = 1, i;
        v[0].resize(N);
        for (int t = 0; t < N; ++t) v[0][t] = r[t] * (1.0 / beta);
        s[0] = beta;
        for (i = 0; i < m && j <= max_iter; i++, j++) {
            matvec(v[i].data(), w.data());
            ++nmv;
            for (int k = 0; k <= i; ++k) {
                double hk = vdot(w, v[k]);
                h(k, i) = hk;
                for (int t = 0; t < N; ++t) w[t] -= hk * v[k][t];
            }
            h(i + 1, i) = vnorm(w);
            v[i + 1].resize(N);
            double inv = 1.0 / h(i + 1, i);
            for (int t = 0; t < N; ++t) v[i + 1][t] = inv * w[t];
            for (int k = 0; k < i; ++k) app_rot(h(k, i), h(k + 1, i), cs[k], sn[k]);
            gen_rot(h(i, i), h(i + 1, i), cs[i], sn[i]);
            app_rot(h(i, i), h(i + 1, i), cs[i], sn[i]);
            app_rot(s[i], s[i + 1], cs[i], sn[i]);
            resid = std::fabs(s[i + 1]) / normb;
            if (resid < tol) {
                update(x, i, h, s, v);
                *resid_out = resid;
                if (matvecs) *matvecs = nmv;
                return j;
            }
        }
        update(x, m - 1, h, s, v);
        matvec(x, w.data());
        ++nmv;
        for (int t = 0; t < N; ++t) r[t] = b[t] - w[t];
        *resid_out = vnorm(r) / normb;
        if (matvecs) *matvecs = nmv;
        return -(j - 1);  // not converged
    }
    template <class HF>
    void update(double *x, int k, HF &h, const Vec &s, const std::vector<Vec> &v) {
        Vec y(s);
        for (int i = k; i >= 0; i--) {
            y[i] /= h(i, i);
            for (int jj = i - 1; jj >= 0; jj--) y[jj] -= h(jj, i) * y[i];
        }
        int N = K.N;
        for (int jj = 0; jj <= k; jj++)
            for (int t = 0; t < N; ++t) x[t] += v[jj][t] * y[jj];
    }
};

int ref_nlevels(int dim, int64_t N, int n_max, int grid) {
    if (grid && dim > 1) {
        int n = (int)std::llround(dim == 2 ? std::sqrt((double)N) : std::cbrt((double)N));
        int p = (int)std::llround(dim == 2 ? std::sqrt((double)n_max) : std::cbrt((double)n_max));
        if (dim == 2) return (int)std::ceil(2.0 * std::log(double(n) / p) / std::log(4));
        return (int)std::ceil(3 * std::log(double(n) / p) / std::log(8));
    }
    return (int)std::ceil(std::log(double(N) / n_max) / std::log((double)(1 << dim)));
}

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

void *orc_create(int dim, int variant, int64_t N, const double *pts, int n_max, double L, int tol_pow, int choice,
                 int nlevels, int compat) {
    if (dim < 1 || dim > 3 || (choice != 0 && choice != 1) || N < 2 || variant < 0 || variant > 2) return nullptr;
    if (variant == 2 && dim == 1) return nullptr;  // the reference has no 1D H* build
    Oracle *o = new Oracle();
    o->dim = dim;
    o->variant = variant;
    o->n_max = n_max;
    o->L = L;
    o->tol_pow = tol_pow;
    o->K.choice = choice;
    int nl = nlevels > 0 ? nlevels : ref_nlevels(dim, N, n_max, pts == nullptr);
    if (nl < 1) {
        delete o;
        return nullptr;
    }
    o->build(pts, N, nl, compat);
    return o;
}
void orc_destroy(void *h) { delete (Oracle *)h; }
int orc_nlevels(void *h) { return ((Oracle *)h)->nL; }
void orc_assemble(void *h) { ((Oracle *)h)->assemble(); }
double orc_memory(void *h) { return ((Oracle *)h)->memory; }
double orc_diag(void *h) { return ((Oracle *)h)->K.diag; }
double orc_kii(void *h) { return ((Oracle *)h)->K.Kii; }
double orc_entry(void *h, int i, int j) { return ((Oracle *)h)->K(i, j); }
void orc_points(void *h, double *out) {
    Oracle *o = (Oracle *)h;
    int N = o->K.N;
    for (int i = 0; i < N; ++i) {
        out[i] = o->K.x[i];
        if (o->dim > 1) out[N + i] = o->K.y[i];
        if (o->dim > 2) out[2 * N + i] = o->K.z[i];
    }
}
void orc_matvec(void *h, const double *x, double *y) { ((Oracle *)h)->matvec(x, y); }
void orc_dense_matvec(void *h, const double *x, double *y) { ((Oracle *)h)->dense_matvec(x, y, nullptr, ((Oracle *)h)->K.N); }
void orc_dense_rows(void *h, const int64_t *rows, int64_t nrows, const double *x, double *y) {
    ((Oracle *)h)->dense_matvec(x, y, rows, nrows);
}
int orc_gmres(void *h, const double *b, double *x, int max_iter, double tol, double *resid, int *matvecs) {
    return ((Oracle *)h)->gmres(b, x, max_iter, tol, resid, matvecs);
}
void orc_gl16(double *nodes, double *weights) { gauss_legendre16(nodes, weights); }
// OpenMP threads a parallel region of the oracle runs on (the CPU baseline's core count)
int orc_threads() { return omp_get_max_threads(); }
// tree order: leaves' point lists concatenated
void orc_tree_order(void *h, int *perm) {
    Oracle *o = (Oracle *)h;
    int c = 0;
    for (auto &B : o->tree[o->nL])
        for (int p : B.pts) perm[c++] = p;
}
int orc_box_count(void *h, int level, int k) { return (int)((Oracle *)h)->tree[level][k].pts.size(); }
// interaction list of box k at a level; vflag receives isVertex
int orc_il(void *h, int level, int k, int *out, int *vflag, int cap) {
    Box &B = ((Oracle *)h)->tree[level][k];
    int n = (int)B.il.size();
    for (int i = 0; i < n && i < cap; ++i) {
        out[i] = B.il[i];
        vflag[i] = B.vert[i];
    }
    return n;
}
// near neighbours in slot order (self excluded)
int orc_near(void *h, int level, int k, int *out, int cap) {
    Oracle *o = (Oracle *)h;
    Box &B = o->tree[level][k];
    int n = 0;
    if (o->dim == 1) return 0;
    for (int s = 0; s < o->nslot; ++s)
        if (B.nbr[s] != -1) {
            if (n < cap) out[n] = B.nbr[s];
            ++n;
        }
    return n;
}
// ranks: family 0 far, 1 vertex (nested)
int orc_rank(void *h, int family, int level, int k) {
    Box &B = ((Oracle *)h)->tree[level][k];
    return family == 0 ? (int)B.Ir.size() : (int)B.Irv.size();
}
// pivots (input-order point indices) of box k: family 0 far (Ir, Ic), 1 vertex
int orc_pivots(void *h, int family, int level, int k, int *ir, int *ic) {
    Box &B = ((Oracle *)h)->tree[level][k];
    const std::vector<int> &a = family == 0 ? B.Ir : B.Irv;
    const std::vector<int> &b = family == 0 ? B.Ic : B.Icv;
    for (size_t i = 0; i < a.size(); ++i) ir[i] = a[i];
    for (size_t i = 0; i < b.size(); ++i) ic[i] = b[i];
    return (int)a.size();
}
// standalone ACA on explicit index sets (input-order indices)
int orc_aca(void *h, const int *rows, int m, const int *cols, int n, int tol_pow, int *rb, int *cb, double *Lout,
            double *Rout) {
    Oracle *o = (Oracle *)h;
    std::vector<int> r(rows, rows + m), c(cols, cols + n);
    AcaResult a;
    aca(o->K, r, c, tol_pow, a, false);
    for (int i = 0; i < a.rank; ++i) {
        rb[i] = a.rb[i];
        cb[i] = a.cb[i];
    }
    if (Lout) std::copy(a.L.a.begin(), a.L.a.end(), Lout);
    if (Rout) std::copy(a.R.a.begin(), a.R.a.end(), Rout);
    return a.rank;
}
}
