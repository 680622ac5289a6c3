// TestDropinHelpers.cpp -- a caller of the NonPeriodic3D drop-in
// (include/hz_compat/NonPeriodic3D_MPI.h) that exercises the interface's
// helper functions one at a time, with the command line of the fixture
// driver's modes of the same names (oracle/ref_driver.cpp), so that a test
// can run both on the same arguments and compare the outputs byte for byte:
//
//   zderiv Nx Ny Nz N zl zr IN OUT        FirstDerivativeZ then
//                                         SecondDerivativeZ of element arrays
//                                         IN [Ns][Nx][Ny][N]; OUT.rank0
//   fundsol Nx Ny Nz N xl xr yl yr zl zr lambda iw jw i j IN OUT
//                                         the test callback's operator of mode
//                                         (iw, jw) through LUDecomposition,
//                                         rhsPDE[kk][i][j] from IN[Ns][N],
//                                         ConstructFundamentalSolutions(U, i, j);
//                                         OUT: U[kk].phi[i][j][0..N)
//   lusolve n COUNT IN OUT                per system A[n*n] b[n]:
//                                         LUDecomposition, LUSolve; OUT: x
//   write Nx Ny Nz N xl xr yl yr zl zr IN WriteFile of U from IN with the
//                                         polynomial exact function below, into
//                                         ./Output/Solution-0000.tec
#include "../../include/hz_compat/NonPeriodic3D_MPI.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static std::vector<double> read_doubles(const char *path, size_t count)
{
    std::vector<double> v(count);
    FILE *fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(double), count, fp) != count) {
        std::fprintf(stderr, "TestDropinHelpers: cannot read %zu doubles from %s\n", count, path);
        std::exit(2);
    }
    std::fclose(fp);
    return v;
}

static void write_raw(FILE *fp, const void *p, size_t bytes)
{
    if (!fp || std::fwrite(p, 1, bytes, fp) != bytes) {
        std::fprintf(stderr, "TestDropinHelpers: write failed\n");
        std::exit(2);
    }
}

// The operator of the reference's test program (TestNonPeriodic3D_MPI.cpp:71-79).
static void construct_operator(const int iw, const int jw)
{
    const double shift = Lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx) + 4.0 * PI * PI * jw * jw / (Ly * Ly);
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dzz[i + 1][j + 1] - (i == j ? shift : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

static double tecplot_exact(const double x, const double y, const double zz)
{
    return x + 2.0 * y - 3.0 * zz + x * y * zz;
}

static void plan_from_args(char **argv, int first)
{
    double b[6];
    for (int i = 0; i < 6; ++i) b[i] = std::atof(argv[first + 4 + i]);
    CreatePDEPlan(std::atoi(argv[first]), std::atoi(argv[first + 1]), std::atoi(argv[first + 2]),
                  std::atoi(argv[first + 3]), b[0], b[1], b[2], b[3], b[4], b[5]);
}

static FieldVariable *new_fields()
{
    FieldVariable *U = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) Allocate(U[kk].phi, 2 * Nx, NY, N);
    return U;
}

static void delete_fields(FieldVariable *U)
{
    for (int kk = 0; kk < Ns; ++kk) Deallocate(U[kk].phi, 2 * Nx, NY, N);
    delete[] U;
}

static void fill_fields(FieldVariable *U, const std::vector<double> &in)
{
    const size_t col = (size_t)Nx * Ny * N;
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(U[kk].phi[i][j], &in[kk * col + ((size_t)i * Ny + j) * N], sizeof(double) * N);
}

static int mode_zderiv(int argc, char **argv)
{
    if (argc < 10) return 2;
    CreatePDEPlan(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), 0.0, 1.0, 0.0, 1.0,
                  std::atof(argv[6]), std::atof(argv[7]));
    FieldVariable *U = new_fields(), *Uz = new_fields(), *Uzz = new_fields();
    fill_fields(U, read_doubles(argv[8], (size_t)Ns * Nx * Ny * N));
    FirstDerivativeZ(Uz, U);
    SecondDerivativeZ(Uzz, U);
    FILE *fp = std::fopen((std::string(argv[9]) + ".rank0").c_str(), "wb");
    for (FieldVariable *F : {Uz, Uzz})
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j) write_raw(fp, F[kk].phi[i][j], sizeof(double) * N);
    std::fclose(fp);
    delete_fields(U);
    delete_fields(Uz);
    delete_fields(Uzz);
    DestroyPDEPlan();
    return 0;
}

static int mode_fundsol(int argc, char **argv)
{
    if (argc < 19) return 2;
    plan_from_args(argv, 2);
    Lambda = std::atof(argv[12]);
    const int iw = std::atoi(argv[13]), jw = std::atoi(argv[14]), i = std::atoi(argv[15]), j = std::atoi(argv[16]);
    std::vector<double> in = read_doubles(argv[17], (size_t)Ns * N);
    FieldVariable *U = new_fields();
    for (int kk = 0; kk < Ns; ++kk)
        for (int k = 0; k < N; ++k) rhsPDE[kk][i][j][k] = in[(size_t)kk * N + k];
    construct_operator(iw, jw);
    ConstructFundamentalSolutions(U, i, j);
    FILE *fp = std::fopen(argv[18], "wb");
    for (int kk = 0; kk < Ns; ++kk) write_raw(fp, U[kk].phi[i][j], sizeof(double) * N);
    std::fclose(fp);
    delete_fields(U);
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
        for (int i = 0; i < n; ++i) std::memcpy(A[i], &in[c * per + (size_t)i * n], sizeof(double) * n);
        std::memcpy(x, &in[c * per + (size_t)n * n], sizeof(double) * n);
        double det = 0.0;
        LUDecomposition(A, piv, det, n);
        LUSolve(A, piv, x, n);
        write_raw(fp, x, sizeof(double) * n);
    }
    std::fclose(fp);
    Deallocate(A, n, n);
    Deallocate(x, n);
    Deallocate(piv, n);
    return 0;
}

static int mode_write(int argc, char **argv)
{
    if (argc < 13) return 2;
    plan_from_args(argv, 2);
    FieldVariable *U = new_fields();
    fill_fields(U, read_doubles(argv[12], (size_t)Ns * Nx * Ny * N));
    WriteFile(U, tecplot_exact);
    delete_fields(U);
    DestroyPDEPlan();
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "zderiv") return mode_zderiv(argc, argv);
    if (mode == "fundsol") return mode_fundsol(argc, argv);
    if (mode == "lusolve") return mode_lusolve(argc, argv);
    if (mode == "write") return mode_write(argc, argv);
    std::fprintf(stderr, "TestDropinHelpers: unknown mode %s\n", mode.c_str());
    return 2;
}
