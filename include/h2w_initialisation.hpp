// h2w_initialisation.hpp -- C++ drop-in for the reference's operator facade and
// GMRES driver, written over the C ABI of libh2w (include/h2w.h).
//
//   h2w::initialisation   replaces  class initialisation      (3D/codes/matvec.hpp:6-178,
//                                                               2D/codes/matvec.hpp:6-172,
//                                                               1D/codes/matvec.hpp:6-113)
//   h2w::iterative_solver replaces  class iterative_solver    (1D/codes/gmres.hpp:14-201)
//
// Same constructor arguments, same call sequence as the reference's main()
// (initialisation(...), tree_initialisation(), GMRES(p, x, b), then resid /
// max_iter), with the format picked at run time instead of by -DUSE_nHODLRdD /
// -DUSE_snHODLRdD.  Vectors are any contiguous double container with data()
// and size() (std::vector<double>, Eigen::VectorXd).  Errors throw
// h2w::error instead of exiting.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "h2w.h"

namespace h2w {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != H2W_OK) throw error(rc, h2w_last_error());
}

class initialisation {
  public:
    int N, nLevels, n_max, TOL_POW, Kernel_Choice, dim, variant;
    double L;
    h2w_handle *K = nullptr;

    // 1D form of the reference constructor: (N, nLevels, n_max, L, TOL_POW, choice)
    // (1D/codes/matvec.hpp:18).  nLevels <= 0 lets the library use the reference
    // main's formula.  points: dim x N SoA, nullptr = the reference's uniform grid.
    initialisation(int dim, int N, int nLevels, int n_max, double L, int TOL_POW, int choice,
                   int variant = H2W_NESTED, const double *points = nullptr, const h2w_opts *opts = nullptr)
        : N(N), nLevels(nLevels), n_max(n_max), TOL_POW(TOL_POW), Kernel_Choice(choice), dim(dim), variant(variant),
          L(L) {
        h2w_opts o;
        h2w_default_opts(&o);
        if (opts) o = *opts;
        if (nLevels > 0) o.nlevels = nLevels;
        check(h2w_create(dim, variant, N, points, n_max, L, TOL_POW, choice, &o, &K));
        h2w_stats st;
        check(h2w_get_stats(K, &st));
        this->nLevels = st.nlevels;
    }
    initialisation(const initialisation &) = delete;
    initialisation &operator=(const initialisation &) = delete;
    ~initialisation() { h2w_destroy(K); }

    // tree_initialisation(): tree, interaction lists, NCA/ACA assembly (3D/codes/matvec.hpp:33-61)
    void tree_initialisation() { check(h2w_assemble(K)); }

    // mat_vec_prod(charge, output): output = A charge, input order (3D/codes/matvec.hpp:63-83)
    template <class V>
    void mat_vec_prod(const V &charge, V &output) {
        output.resize(charge.size());
        check(h2w_matvec(K, charge.data(), output.data(), H2W_HOST));
    }

    // the reference's storage counter (findMemory_n / findMemory_sn)
    double memory() const {
        h2w_stats st;
        check(h2w_get_stats(K, &st));
        return st.memory;
    }

    // getTruePoten: dense FP64 product, on the GPU (3D/codes/HODLR3Dtree.hpp:2630-2643)
    template <class V>
    V getTruePoten(const V &charge) const {
        V out(charge.size());
        check(h2w_dense_matvec(K, charge.data(), out.data(), nullptr, 0, H2W_HOST));
        return out;
    }
};

class iterative_solver {
  public:
    int restart, max_iter;
    double gmres_tol, resid;
    int mode;
    h2w_gmres_stats stats{};

    // (max_iter, gmres_tol_pow) as 1D/codes/gmres.hpp:22-28; mode selects FP64 or FP32 + refinement
    iterative_solver(int max_iter, int gmres_tol_pow, int mode = H2W_GMRES_FP64)
        : restart(max_iter), max_iter(max_iter), gmres_tol(std::pow(10, -1.0 * gmres_tol_pow)), resid(0.0),
          mode(mode) {}

    // GMRES(A, x, b): x holds x0 on entry, the solution on return; max_iter then holds the
    // iteration count and resid the final residual, as in the reference (1D/codes/gmres.hpp:162-170)
    template <class T, class V>
    void GMRES(T *&A, V &x, const V &b) {
        int rc = h2w_gmres(A->K, b.data(), x.data(), max_iter, gmres_tol, mode, H2W_HOST, &stats);
        if (rc != H2W_OK && rc != H2W_ENOCONV) check(rc);
        resid = stats.resid;
        max_iter = stats.iterations;
    }
};

}  // namespace h2w
