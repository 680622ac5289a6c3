// ref_driver_periodic.cpp -- fixture driver for gpritam/MPI's periodic-z
// DirectPseudoSpectral solver (HelmholtzSolvers/DirectPseudoSpectral/src/
// Periodic3D_MPI.cpp), compiled where the reference sources lie by
// oracle/Makefile into oracle/_ref/ (never committed).
//
// TEST INFRASTRUCTURE ONLY: it writes golden vectors under tests/golden/.
// It contains no reference code; it only calls the reference's public
// functions (include/Periodic3D_MPI.h).  A separate binary from ref_driver
// because both solvers define the same global names.
//
//   psolve Nx Ny Nz N xl xr yl yr zl zr lambda IN OUT
//       (under mpiexec -np P) IN: f[Nx][Ny][Ktot]; OUT.rankR: this rank's
//       [Ns][Nx][Ny][N] block, then z; OUT.alpha (rank 0): the closure
//       coefficients alphaPDE[Nx][2 Ny] after the solve (broadcast to every
//       rank by the reference; zeros when the run has one element).
#include "Periodic3D_MPI.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static void write_raw(FILE *fp, const void *p, size_t bytes)
{
    if (std::fwrite(p, 1, bytes, fp) != bytes) {
        std::fprintf(stderr, "ref_driver_periodic: write failed\n");
        std::exit(2);
    }
}

// The operator callback of the reference's test program
// (TestPeriodic3D_MPI.cpp semantics): interior Dzz minus the Fourier shift,
// then the reference's LU factorisation.
static void construct_operator(const int iw, const int jw)
{
    const double shift = Lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx) + 4.0 * PI * PI * jw * jw / (Ly * Ly);
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dzz[i + 1][j + 1] - (i == j ? shift : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

int main(int argc, char **argv)
{
    if (argc < 15 || std::string(argv[1]) != "psolve") {
        std::fprintf(stderr, "usage: psolve Nx Ny Nz N xl xr yl yr zl zr lambda IN OUT\n");
        return 2;
    }
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &Rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]), np = std::atoi(argv[5]);
    double b[6];
    for (int i = 0; i < 6; ++i) b[i] = std::atof(argv[6 + i]);
    CreatePDEPlan(nx, ny, nz, np, b[0], b[1], b[2], b[3], b[4], b[5]);
    CreateFFTPlan(Maximum(Nx, Ny));
    Lambda = std::atof(argv[12]);
    const int ktot = size * Ns * (N - 1) + 1;
    const size_t nf = (size_t)Nx * Ny * ktot;
    std::vector<double> in(nf);
    FILE *fi = std::fopen(argv[13], "rb");
    if (!fi || std::fread(in.data(), sizeof(double), nf, fi) != nf) {
        std::fprintf(stderr, "ref_driver_periodic: cannot read %s\n", argv[13]);
        return 2;
    }
    std::fclose(fi);
    FieldVariable *V = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) Allocate(V[kk].phi, 2 * Nx, NY, N);
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    rhsPDE[kk][i][j][k] = in[((size_t)i * Ny + j) * ktot + (size_t)(Rank * Ns + kk) * (N - 1) + k];
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    SolvePDE(V, construct_operator);
    MPI_Barrier(MPI_COMM_WORLD);
    const double t1 = MPI_Wtime();
    std::string path = std::string(argv[14]) + ".rank" + std::to_string(Rank);
    FILE *fp = std::fopen(path.c_str(), "wb");
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) write_raw(fp, V[kk].phi[i][j], sizeof(double) * N);
    write_raw(fp, z, sizeof(double) * (Ns * (N - 1) + 1));
    std::fclose(fp);
    if (Rank == 0) {
        FILE *fa = std::fopen((std::string(argv[14]) + ".alpha").c_str(), "wb");
        for (int i = 0; i < Nx; ++i) write_raw(fa, alphaPDE[i], sizeof(double) * NY);
        std::fclose(fa);
        std::printf("solve_seconds %.6f Ns %d size %d\n", t1 - t0, Ns, size);
    }
    for (int kk = 0; kk < Ns; ++kk) Deallocate(V[kk].phi, 2 * Nx, NY, N);
    delete[] V;
    DestroyFFTPlan(Maximum(Nx, Ny));
    DestroyPDEPlan();
    MPI_Finalize();
    return 0;
}
