// ref_driver_low.cpp -- fixture driver for gpritam/MPI's 2-D and 1-D
// DirectPseudoSpectral solvers (HelmholtzSolvers/DirectPseudoSpectral/src/
// NonPeriodic2D_MPI.cpp, NonPeriodic1D_MPI.cpp, Periodic2D_MPI.cpp,
// Periodic1D_MPI.cpp), compiled where the reference sources lie by
// oracle/Makefile into oracle/_ref/ (never committed, never shipped), once
// per solver: HZ_DIM = 2 or 1, HZ_PERIODIC = 0 or 1.
//
// TEST INFRASTRUCTURE ONLY: it writes golden vectors under tests/golden/.
// It contains no reference code; it only calls the reference's public
// functions (the solvers' headers under include/).
//
//   2-D: solve Nx Ny N xl xr yl yr al bl ar br lambda IN OUT
//   1-D: solve Nx N xl xr al bl ar br lambda IN OUT
// (under mpiexec -np P; the Robin coefficients are ignored by the periodic
// solvers).  Ktot = P*Ns*(N-1)+1 grid points along the GLL axis.
//   IN : 2-D f[Nx][Ktot], then (non-periodic) gl[Nx], gr[Nx];
//        1-D f[Ktot], then (non-periodic) gl, gr.
//   OUT.rankR: this rank's block, [Ns][Nx][N] (2-D) or [Ns][N] (1-D), then
//        its grid (Ns*(N-1)+1 points).
#if HZ_DIM == 2
#if HZ_PERIODIC
#include "Periodic2D_MPI.h"
#else
#include "NonPeriodic2D_MPI.h"
#endif
#else
#if HZ_PERIODIC
#include "Periodic1D_MPI.h"
#else
#include "NonPeriodic1D_MPI.h"
#endif
#endif

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static void write_raw(FILE *fp, const void *p, size_t bytes)
{
    if (std::fwrite(p, 1, bytes, fp) != bytes) {
        std::fprintf(stderr, "ref_driver_low: write failed\n");
        std::exit(2);
    }
}

// The operator callbacks of the reference's test programs: the interior
// second-derivative matrix minus the (shifted) Helmholtz constant, then the
// reference's LU factorisation.
#if HZ_DIM == 2
static void construct_operator(const int iw)
{
    const double shift = Lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx);
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dyy[i + 1][j + 1] - (i == j ? shift : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}
#else
static void construct_operator()
{
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dxx[i + 1][j + 1] - (i == j ? Lambda : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}
#endif

int main(int argc, char **argv)
{
    const int need = HZ_DIM == 2 ? 15 : 11;
    if (argc < need || std::string(argv[1]) != "solve") {
        std::fprintf(stderr, HZ_DIM == 2 ? "usage: solve Nx Ny N xl xr yl yr al bl ar br lambda IN OUT\n"
                                         : "usage: solve Nx N xl xr al bl ar br lambda IN OUT\n");
        return 2;
    }
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &Rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    int ai = 2;
    const int nx = std::atoi(argv[ai++]);
#if HZ_DIM == 2
    const int ny = std::atoi(argv[ai++]);
#endif
    const int np = std::atoi(argv[ai++]);
    const double a_xl = std::atof(argv[ai++]), a_xr = std::atof(argv[ai++]);
#if HZ_DIM == 2
    const double a_yl = std::atof(argv[ai++]), a_yr = std::atof(argv[ai++]);
    CreatePDEPlan(nx, ny, np, a_xl, a_xr, a_yl, a_yr);
    CreateFFTPlan(Nx);
    const int lines = Nx;
    const double *grid = y;
#else
    CreatePDEPlan(nx, np, a_xl, a_xr);
    const int lines = 1;
    const double *grid = x;
#endif
    double robin[4];
    for (int i = 0; i < 4; ++i) robin[i] = std::atof(argv[ai++]);
#if !HZ_PERIODIC
    alphal = robin[0];
    betal = robin[1];
    alphar = robin[2];
    betar = robin[3];
#endif
    Lambda = std::atof(argv[ai++]);
    const char *in_path = argv[ai++], *out_path = argv[ai++];
    const int ktot = size * Ns * (N - 1) + 1;
    const size_t nf = (size_t)lines * ktot + (HZ_PERIODIC ? 0 : 2 * (size_t)lines);
    std::vector<double> in(nf);
    FILE *fi = std::fopen(in_path, "rb");
    if (!fi || std::fread(in.data(), sizeof(double), nf, fi) != nf) {
        std::fprintf(stderr, "ref_driver_low: cannot read %s\n", in_path);
        return 2;
    }
    std::fclose(fi);
    FieldVariable *V = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) {
#if HZ_DIM == 2
        Allocate(V[kk].phi, NX, N);
        for (int i = 0; i < Nx; ++i)
            for (int k = 0; k < N; ++k)
                rhsPDE[kk][i][k] = in[(size_t)i * ktot + (size_t)(Rank * Ns + kk) * (N - 1) + k];
#else
        Allocate(V[kk].phi, N);
        for (int k = 0; k < N; ++k) rhsPDE[kk][k] = in[(size_t)(Rank * Ns + kk) * (N - 1) + k];
#endif
    }
#if !HZ_PERIODIC
    const double *faces = in.data() + (size_t)lines * ktot;
#if HZ_DIM == 2
    for (int i = 0; i < Nx; ++i) {
        leftBC[i] = faces[i];
        rightBC[i] = faces[Nx + i];
    }
#else
    leftBC = faces[0];
    rightBC = faces[1];
#endif
#endif
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    SolvePDE(V, construct_operator);
    MPI_Barrier(MPI_COMM_WORLD);
    const double t1 = MPI_Wtime();
    std::string path = std::string(out_path) + ".rank" + std::to_string(Rank);
    FILE *fp = std::fopen(path.c_str(), "wb");
    for (int kk = 0; kk < Ns; ++kk) {
#if HZ_DIM == 2
        for (int i = 0; i < Nx; ++i) write_raw(fp, V[kk].phi[i], sizeof(double) * N);
#else
        write_raw(fp, V[kk].phi, sizeof(double) * N);
#endif
    }
    write_raw(fp, grid, sizeof(double) * (Ns * (N - 1) + 1));
    std::fclose(fp);
    if (Rank == 0) std::printf("solve_seconds %.6f Ns %d size %d\n", t1 - t0, Ns, size);
    for (int kk = 0; kk < Ns; ++kk) {
#if HZ_DIM == 2
        Deallocate(V[kk].phi, NX, N);
#else
        Deallocate(V[kk].phi, N);
#endif
    }
    delete[] V;
#if HZ_DIM == 2
    DestroyFFTPlan(Nx);
#endif
    DestroyPDEPlan();
    MPI_Finalize();
    return 0;
}
