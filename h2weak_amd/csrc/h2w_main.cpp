// h2w_main.cpp -- command-line driver over include/h2w_initialisation.hpp, with
// the reference main()'s arguments, call sequence and summary
// (1D/codes/main.cpp:3-82, 2D/codes/main.cpp, 3D/codes/main.cpp):
//
//   h2w_main dim a1 a2 L tol_pow choice [variant] [mode] [seed]
//     dim 1: a1 = N,       a2 = n_max
//     dim 2: a1 = sqrt(N), a2 = leaf points per axis      (n_max = a2^2)
//     dim 3: a1 = cbrt(N), a2 = leaf points per axis      (n_max = a2^3)
//   variant 0 = H2* (nested), 1 = (H2+H)* (semi-nested); mode = h2w GMRES mode.
// The right-hand side is b = A x_true with x_true ~ U(-1,1) from a seeded
// generator (the reference uses Eigen's unseeded Vec::Random), x0 = 1, and the
// GMRES tolerance / max_iter of the reference (1e-12 / 500 in 1D and 2D,
// 1e-10 / 1000 in 3D).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/h2w_initialisation.hpp"

int main(int argc, char **argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s dim a1 a2 L tol_pow choice [variant] [mode] [seed]\n", argv[0]);
        return 2;
    }
    int dim = std::atoi(argv[1]);
    int a1 = std::atoi(argv[2]), a2 = std::atoi(argv[3]);
    double L = std::atof(argv[4]);
    int TOL_POW = std::atoi(argv[5]);
    int Choice = std::atoi(argv[6]);
    int variant = argc > 7 ? std::atoi(argv[7]) : H2W_NESTED;
    int mode = argc > 8 ? std::atoi(argv[8]) : H2W_GMRES_FP64;
    unsigned seed = argc > 9 ? (unsigned)std::atoi(argv[9]) : 2u;
    int N = dim == 1 ? a1 : (dim == 2 ? a1 * a1 : a1 * a1 * a1);
    int n_max = dim == 1 ? a2 : (dim == 2 ? a2 * a2 : a2 * a2 * a2);
    try {
        auto t0 = std::chrono::steady_clock::now();
        h2w::initialisation *p = new h2w::initialisation(dim, N, 0, n_max, L, TOL_POW, Choice, variant);
        p->tree_initialisation();
        double elapsed_assem = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        std::mt19937_64 gen(seed);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        std::vector<double> b(N);
        for (auto &v : b) v = U(gen);
        std::vector<double> Ab_true = p->getTruePoten(b);
        std::vector<double> x(N, 1.0);

        int TOL_GMRES = dim == 3 ? 10 : 12;
        int MAX_ITER = dim == 3 ? 1000 : 500;
        t0 = std::chrono::steady_clock::now();
        h2w::iterative_solver *its = new h2w::iterative_solver(MAX_ITER, TOL_GMRES, mode);
        its->GMRES(p, x, Ab_true);
        double elapsed_gmres = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        double num = 0.0, den = 0.0;
        for (int i = 0; i < N; ++i) {
            num += (x[i] - b[i]) * (x[i] - b[i]);
            den += b[i] * b[i];
        }
        double memory = p->memory();
        std::printf("********** Summary of %s accelerated GMRES (%dD, libh2w on the GPU) **********\n\n",
                    variant == H2W_NESTED ? "H2* (nested)" : "(H2+H)* (semi-nested)", dim);
        std::printf("The number of particles taken: %d and choice =  %s\n\n", N,
                    Choice == 0 ? "integral equation" : "RBF interpolation");
        std::printf("The maximum number of particles at leaf clusters: %d\n\n", n_max);
        std::printf("Depth of the tree: %d\n\n", p->nLevels);
        std::printf("The final residual error is: %g\n\n", its->resid);
        std::printf("The total number of iteration is: %d\n\n", its->max_iter);
        std::printf("Total Assembly time: %gs\n\n", elapsed_assem);
        std::printf("GMRES time: %gs\n\n", elapsed_gmres);
        std::printf("Storage (in GB): %g GB\n\n", memory * 8 * std::pow(10.0, -9));
        std::printf("Compression ratio: %g\n\n", memory / (1.0 * N * N));
        std::printf("The (norm-2) relative error in solution: %g\n", std::sqrt(num / den));
        delete its;
        delete p;
    } catch (const h2w::error &e) {
        std::fprintf(stderr, "h2w error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
