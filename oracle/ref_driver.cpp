// ref_driver.cpp -- fixture driver linked against gpritam/MPI's
// DirectPseudoSpectral sources, compiled where they lie under the reference
// tree by oracle/Makefile into oracle/_ref/ (never committed).  bench.py's
// CPU baseline runs it on the GPU machine's host cores, which is why the
// built binary travels with the repository snapshot there.
//
// TEST INFRASTRUCTURE ONLY: it writes the golden vectors under tests/golden/
// that pin the oracle (hz_oracle.c).  It contains no reference code; it only
// calls the reference's public functions (headers under
// HelmholtzSolvers/DirectPseudoSpectral/include).
//
// Modes (all binary files are raw little-endian float64 / int32):
//   gll N OUT                       x[N] w[N] D[N*N] Dzz[N*N]
//   lu n COUNT IN OUT               IN: COUNT*n*n doubles; OUT per matrix:
//                                   LU[n*n] det, then COUNT*n int32 pivots
//   fft Nx Ny Nz ISIGN IN OUT       FourierTransformXY on A[2Nx][2Ny][Nz]
//   fderiv Nx Ny Nz Lx Ly IN OUT    FirstDerivativeFourierX/Y, SecondDerivative-
//                                   FourierX/Y on F[Nx][Ny][Nz]; OUT: the four
//                                   results in that order
//   solve Nx Ny Nz N xl xr yl yr zl zr al bl ar br lambda IN OUT [REPS]
//                                   (run under mpiexec -np P) IN: f[Nx][Ny][Ktot]
//                                   gl[Nx][Ny] gr[Nx][Ny]; OUT.rankR: this
//                                   rank's [Ns][Nx][Ny][N] block, then z.
//                                   REPS solves (default 1), the inputs
//                                   refilled before each (SolvePDE destroys
//                                   them); rank 0 prints every solve's
//                                   barrier-to-barrier seconds.
//   zderiv Nx Ny Nz N zl zr IN OUT  (under mpiexec -np P) IN: element arrays
//                                   u[P*Ns][Nx][Ny][N]; OUT.rankR: this rank's
//                                   FirstDerivativeZ then SecondDerivativeZ
//                                   blocks [Ns][Nx][Ny][N]
//   fundsol Nx Ny Nz N xl xr yl yr zl zr lambda iw jw i j IN OUT
//                                   operator of mode (iw,jw) factored by the
//                                   test callback, rhsPDE[kk][i][j][0..N) from
//                                   IN[Ns][N], then ConstructFundamentalSolutions
//                                   (U, i, j); OUT: U[kk].phi[i][j][0..N)
//   lusolve n COUNT IN OUT          IN per system: A[n*n] b[n]; OUT per system:
//                                   LUSolve's x[n] after LUDecomposition
//   write Nx Ny Nz N xl xr yl yr zl zr IN
//                                   U[Ns][Nx][Ny][N] from IN, then WriteFile
//                                   with tecplot_exact (below) into
//                                   ./Output/Solution-0000.tec
#include "NonPeriodic3D_MPI.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static std::vector<double> read_doubles(const char *path, size_t count)
{
    std::vector<double> v(count);
    FILE *fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(double), count, fp) != count) {
        std::fprintf(stderr, "ref_driver: cannot read %zu doubles from %s\n", count, path);
        std::exit(2);
    }
    std::fclose(fp);
    return v;
}

static void write_raw(FILE *fp, const void *p, size_t bytes)
{
    if (std::fwrite(p, 1, bytes, fp) != bytes) {
        std::fprintf(stderr, "ref_driver: write failed\n");
        std::exit(2);
    }
}

static int mode_gll(int argc, char **argv)
{
    if (argc < 4) return 2;
    const int n = std::atoi(argv[2]);
    double *x, *w, **D, **Dzz;
    Allocate(x, n);
    Allocate(w, n);
    Allocate(D, n, n);
    Allocate(Dzz, n, n);
    GaussLobattoLegendrePoints(x, n);
    GaussLobattoLegendreWeights(w, x, n);
    FirstDerivativeMatrix(D, x, n);
    SecondDerivativeMatrix(Dzz, D, n);
    FILE *fp = std::fopen(argv[3], "wb");
    write_raw(fp, x, sizeof(double) * n);
    write_raw(fp, w, sizeof(double) * n);
    for (int i = 0; i < n; ++i) write_raw(fp, D[i], sizeof(double) * n);
    for (int i = 0; i < n; ++i) write_raw(fp, Dzz[i], sizeof(double) * n);
    std::fclose(fp);
    return 0;
}

static int mode_lu(int argc, char **argv)
{
    if (argc < 6) return 2;
    const int n = std::atoi(argv[2]), count = std::atoi(argv[3]);
    std::vector<double> in = read_doubles(argv[4], (size_t)count * n * n);
    FILE *fp = std::fopen(argv[5], "wb");
    double **A;
    int *piv;
    Allocate(A, n, n);
    Allocate(piv, n);
    std::vector<int> pivots;
    for (int c = 0; c < count; ++c) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) A[i][j] = in[((size_t)c * n + i) * n + j];
        double det = 0.0;
        LUDecomposition(A, piv, det, n);
        for (int i = 0; i < n; ++i) write_raw(fp, A[i], sizeof(double) * n);
        write_raw(fp, &det, sizeof(double));
        for (int i = 0; i < n; ++i) pivots.push_back(i < n - 1 ? piv[i] : -1);
    }
    write_raw(fp, pivots.data(), sizeof(int) * pivots.size());
    std::fclose(fp);
    return 0;
}

static int mode_fft(int argc, char **argv)
{
    if (argc < 8) return 2;
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
    const int isign = std::atoi(argv[5]);
    std::vector<double> in = read_doubles(argv[6], (size_t)4 * nx * ny * nz);
    CreateFFTPlan(nx > ny ? nx : ny);
    FILE *fp = std::fopen(argv[7], "wb");
    if (nz == 1) {
        double **A;
        Allocate(A, 2 * nx, 2 * ny);
        for (int i = 0; i < 2 * nx; ++i)
            for (int j = 0; j < 2 * ny; ++j) A[i][j] = in[(size_t)i * 2 * ny + j];
        FourierTransformXY(A, isign, nx, ny);
        for (int i = 0; i < 2 * nx; ++i) write_raw(fp, A[i], sizeof(double) * 2 * ny);
    } else {
        double ***A;
        Allocate(A, 2 * nx, 2 * ny, nz);
        for (int i = 0; i < 2 * nx; ++i)
            for (int j = 0; j < 2 * ny; ++j)
                for (int k = 0; k < nz; ++k) A[i][j][k] = in[((size_t)i * 2 * ny + j) * nz + k];
        FourierTransformXY(A, isign, nx, ny, nz);
        for (int i = 0; i < 2 * nx; ++i)
            for (int j = 0; j < 2 * ny; ++j) write_raw(fp, A[i][j], sizeof(double) * nz);
    }
    std::fclose(fp);
    return 0;
}

// The operator callback of the reference's own test program: interior rows
// and columns of the scaled Dzz minus the Fourier shift, then the
// reference's LU factorisation (TestNonPeriodic3D_MPI.cpp:71-79 semantics).
static void construct_operator(const int iw, const int jw)
{
    const double shift = Lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx) + 4.0 * PI * PI * jw * jw / (Ly * Ly);
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dzz[i + 1][j + 1] - (i == j ? shift : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

static int mode_solve(int argc, char **argv)
{
    if (argc < 19) return 2;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &Rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]), np = std::atoi(argv[5]);
    double b[10];
    for (int i = 0; i < 6; ++i) b[i] = std::atof(argv[6 + i]);
    CreatePDEPlan(nx, ny, nz, np, b[0], b[1], b[2], b[3], b[4], b[5]);
    CreateFFTPlan(Maximum(Nx, Ny));
    alphal = std::atof(argv[12]);
    betal = std::atof(argv[13]);
    alphar = std::atof(argv[14]);
    betar = std::atof(argv[15]);
    Lambda = std::atof(argv[16]);
    const int ktot = size * Ns * (N - 1) + 1;
    const size_t nf = (size_t)Nx * Ny * ktot, np2 = (size_t)Nx * Ny;
    std::vector<double> in = read_doubles(argv[17], nf + 2 * np2);
    FieldVariable *V = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) Allocate(V[kk].phi, 2 * Nx, NY, N);
    const int reps = argc > 19 ? std::atoi(argv[19]) : 1;
    std::vector<double> secs;
    for (int rep = 0; rep < (reps > 0 ? reps : 1); ++rep) {
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < N; ++k)
                        rhsPDE[kk][i][j][k] = in[((size_t)i * Ny + j) * ktot + (size_t)(Rank * Ns + kk) * (N - 1) + k];
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) {
                leftBC[i][j] = in[nf + (size_t)i * Ny + j];
                rightBC[i][j] = in[nf + np2 + (size_t)i * Ny + j];
            }
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        SolvePDE(V, construct_operator);
        MPI_Barrier(MPI_COMM_WORLD);
        secs.push_back(MPI_Wtime() - t0);
    }
    std::string path = std::string(argv[18]) + ".rank" + std::to_string(Rank);
    FILE *fp = std::fopen(path.c_str(), "wb");
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) write_raw(fp, V[kk].phi[i][j], sizeof(double) * N);
    write_raw(fp, z, sizeof(double) * (Ns * (N - 1) + 1));
    std::fclose(fp);
    if (Rank == 0) {
        std::printf("solve_seconds");
        for (double t : secs) std::printf(" %.6f", t);
        std::printf(" Ns %d size %d\n", Ns, size);
    }
    for (int kk = 0; kk < Ns; ++kk) Deallocate(V[kk].phi, 2 * Nx, NY, N);
    delete[] V;
    DestroyFFTPlan(Maximum(Nx, Ny));
    DestroyPDEPlan();
    MPI_Finalize();
    return 0;
}

static int mode_fderiv(int argc, char **argv)
{
    if (argc < 9) return 2;
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
    const double lx = std::atof(argv[5]), ly = std::atof(argv[6]);
    const size_t count = (size_t)nx * ny * nz;
    std::vector<double> in = read_doubles(argv[7], count);
    CreateFFTPlan(Maximum(nx, ny));
    double ***F, ***G;
    Allocate(F, nx, ny, nz);
    Allocate(G, nx, ny, nz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) F[i][j][k] = in[((size_t)i * ny + j) * nz + k];
    FILE *fp = std::fopen(argv[8], "wb");
    for (int which = 0; which < 4; ++which) {
        if (which == 0) FirstDerivativeFourierX(G, F, nx, ny, nz, lx);
        if (which == 1) FirstDerivativeFourierY(G, F, nx, ny, nz, ly);
        if (which == 2) SecondDerivativeFourierX(G, F, nx, ny, nz, lx);
        if (which == 3) SecondDerivativeFourierY(G, F, nx, ny, nz, ly);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) write_raw(fp, G[i][j], sizeof(double) * nz);
    }
    std::fclose(fp);
    Deallocate(F, nx, ny, nz);
    Deallocate(G, nx, ny, nz);
    DestroyFFTPlan(Maximum(nx, ny));
    return 0;
}

static int mode_zderiv(int argc, char **argv)
{
    if (argc < 10) return 2;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &Rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]), np = std::atoi(argv[5]);
    CreatePDEPlan(nx, ny, nz, np, 0.0, 1.0, 0.0, 1.0, std::atof(argv[6]), std::atof(argv[7]));
    const size_t col = (size_t)Nx * Ny * N;
    std::vector<double> in = read_doubles(argv[8], (size_t)size * Ns * col);
    FieldVariable *U = new FieldVariable[Ns], *Uz = new FieldVariable[Ns], *Uzz = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) {
        Allocate(U[kk].phi, 2 * Nx, NY, N);
        Allocate(Uz[kk].phi, 2 * Nx, NY, N);
        Allocate(Uzz[kk].phi, 2 * Nx, NY, N);
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    U[kk].phi[i][j][k] = in[(size_t)(Rank * Ns + kk) * col + ((size_t)i * Ny + j) * N + k];
    }
    FirstDerivativeZ(Uz, U);
    SecondDerivativeZ(Uzz, U);
    std::string path = std::string(argv[9]) + ".rank" + std::to_string(Rank);
    FILE *fp = std::fopen(path.c_str(), "wb");
    for (FieldVariable *F : {Uz, Uzz})
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j) write_raw(fp, F[kk].phi[i][j], sizeof(double) * N);
    std::fclose(fp);
    for (int kk = 0; kk < Ns; ++kk) {
        Deallocate(U[kk].phi, 2 * Nx, NY, N);
        Deallocate(Uz[kk].phi, 2 * Nx, NY, N);
        Deallocate(Uzz[kk].phi, 2 * Nx, NY, N);
    }
    delete[] U;
    delete[] Uz;
    delete[] Uzz;
    DestroyPDEPlan();
    MPI_Finalize();
    return 0;
}

static int mode_fundsol(int argc, char **argv)
{
    if (argc < 19) return 2;
    Rank = 0;
    size = 1;
    double b[6];
    for (int i = 0; i < 6; ++i) b[i] = std::atof(argv[6 + i]);
    CreatePDEPlan(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), b[0], b[1], b[2],
                  b[3], b[4], b[5]);
    Lambda = std::atof(argv[12]);
    const int iw = std::atoi(argv[13]), jw = std::atoi(argv[14]), i = std::atoi(argv[15]), j = std::atoi(argv[16]);
    std::vector<double> in = read_doubles(argv[17], (size_t)Ns * N);
    FieldVariable *U = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) {
        Allocate(U[kk].phi, 2 * Nx, NY, N);
        for (int k = 0; k < N; ++k) rhsPDE[kk][i][j][k] = in[(size_t)kk * N + k];
    }
    construct_operator(iw, jw);
    ConstructFundamentalSolutions(U, i, j);
    FILE *fp = std::fopen(argv[18], "wb");
    for (int kk = 0; kk < Ns; ++kk) write_raw(fp, U[kk].phi[i][j], sizeof(double) * N);
    std::fclose(fp);
    for (int kk = 0; kk < Ns; ++kk) Deallocate(U[kk].phi, 2 * Nx, NY, N);
    delete[] U;
    DestroyPDEPlan();
    return 0;
}

static int mode_lusolve(int argc, char **argv)
{
    if (argc < 6) return 2;
    const int n = std::atoi(argv[2]), count = std::atoi(argv[3]);
    const size_t per = (size_t)n * n + n;
    std::vector<double> in = read_doubles(argv[4], (size_t)count * per);
    FILE *fp = std::fopen(argv[5], "wb");
    double **A, *x;
    int *piv;
    Allocate(A, n, n);
    Allocate(x, n);
    Allocate(piv, n);
    for (int c = 0; c < count; ++c) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) A[i][j] = in[c * per + (size_t)i * n + j];
        for (int i = 0; i < n; ++i) x[i] = in[c * per + (size_t)n * n + i];
        double det = 0.0;
        LUDecomposition(A, piv, det, n);
        LUSolve(A, piv, x, n);
        write_raw(fp, x, sizeof(double) * n);
    }
    std::fclose(fp);
    return 0;
}

// The exact function handed to WriteFile: a polynomial, so that the
// reference build and a drop-in caller evaluate it to the same bits.
static double tecplot_exact(const double x, const double y, const double zz)
{
    return x + 2.0 * y - 3.0 * zz + x * y * zz;
}

static int mode_write(int argc, char **argv)
{
    if (argc < 13) return 2;
    Rank = 0;
    size = 1;
    double b[6];
    for (int i = 0; i < 6; ++i) b[i] = std::atof(argv[6 + i]);
    CreatePDEPlan(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), b[0], b[1], b[2],
                  b[3], b[4], b[5]);
    const size_t col = (size_t)Nx * Ny * N;
    std::vector<double> in = read_doubles(argv[12], (size_t)Ns * col);
    FieldVariable *U = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) {
        Allocate(U[kk].phi, 2 * Nx, NY, N);
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k) U[kk].phi[i][j][k] = in[(size_t)kk * col + ((size_t)i * Ny + j) * N + k];
    }
    WriteFile(U, tecplot_exact);
    for (int kk = 0; kk < Ns; ++kk) Deallocate(U[kk].phi, 2 * Nx, NY, N);
    delete[] U;
    DestroyPDEPlan();
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "gll") return mode_gll(argc, argv);
    if (mode == "lu") return mode_lu(argc, argv);
    if (mode == "fft") return mode_fft(argc, argv);
    if (mode == "solve") return mode_solve(argc, argv);
    if (mode == "fderiv") return mode_fderiv(argc, argv);
    if (mode == "zderiv") return mode_zderiv(argc, argv);
    if (mode == "fundsol") return mode_fundsol(argc, argv);
    if (mode == "lusolve") return mode_lusolve(argc, argv);
    if (mode == "write") return mode_write(argc, argv);
    std::fprintf(stderr, "ref_driver: unknown mode %s\n", mode.c_str());
    return 2;
}
