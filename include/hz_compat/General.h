// General.h (drop-in) -- the subset of gpritam/MPI's
// HelmholtzSolvers/DirectPseudoSpectral/include/General.h that callers of the
// DirectPseudoSpectral solver use: contiguous zero-filled N-d arrays with
// row pointer tables (General.h:305-684), small helpers and constants
// (General.h:12-15, :1091-1103) and ErrorMessage (General.cpp:6-11).
// Same names and semantics, independent implementation.
#pragma once

#include <cmath>
#include <cstdlib>
#include <iostream>

const double PI = 3.1415926535897932;
const double EPSILON = 1.000000000000E-12;

void ErrorMessage(const char *message);  // prints, then exit(1)

template <class T> static T Absolute(T m) { return m >= 0.0 ? m : -m; }
template <class T> static T Maximum(T a, T b) { return a > b ? a : b; }
template <class T> static T Minimum(T a, T b) { return a < b ? a : b; }
template <class T> static T Sign(T a, T b) { return b >= 0.0 ? Absolute(a) : -Absolute(a); }

namespace hz_compat_detail {
template <class T> T *block(size_t n)
{
    T *p = static_cast<T *>(std::calloc(n ? n : 1, sizeof(T)));
    if (!p) ErrorMessage("Error: Memory can not be allocated!");
    return p;
}
}  // namespace hz_compat_detail

// One contiguous, zero-filled block per array; higher dimensions are tables
// of pointers into it, so &A[0][0]... is the flat buffer.
template <class T> static void Allocate(T *&A, const int d0, const int = 0)
{
    A = hz_compat_detail::block<T>((size_t)d0);
}
template <class T> static void Allocate(T **&A, const int d0, const int d1, const int = 0, const int = 0)
{
    T *data = hz_compat_detail::block<T>((size_t)d0 * d1);
    A = hz_compat_detail::block<T *>((size_t)d0);
    for (int i = 0; i < d0; ++i) A[i] = data + (size_t)i * d1;
}
template <class T>
static void Allocate(T ***&A, const int d0, const int d1, const int d2, const int = 0, const int = 0, const int = 0)
{
    T *data = hz_compat_detail::block<T>((size_t)d0 * d1 * d2);
    T **rows = hz_compat_detail::block<T *>((size_t)d0 * d1);
    A = hz_compat_detail::block<T **>((size_t)d0);
    for (int i = 0; i < d0; ++i) {
        A[i] = rows + (size_t)i * d1;
        for (int j = 0; j < d1; ++j) A[i][j] = data + ((size_t)i * d1 + j) * d2;
    }
}
template <class T>
static void Allocate(T ****&A, const int d0, const int d1, const int d2, const int d3, const int = 0,
                     const int = 0, const int = 0, const int = 0)
{
    A = hz_compat_detail::block<T ***>((size_t)d0);
    T *data = hz_compat_detail::block<T>((size_t)d0 * d1 * d2 * d3);
    T **rows = hz_compat_detail::block<T *>((size_t)d0 * d1 * d2);
    T ***planes = hz_compat_detail::block<T **>((size_t)d0 * d1);
    for (int i = 0; i < d0; ++i) {
        A[i] = planes + (size_t)i * d1;
        for (int j = 0; j < d1; ++j) {
            A[i][j] = rows + ((size_t)i * d1 + j) * d2;
            for (int k = 0; k < d2; ++k) A[i][j][k] = data + (((size_t)i * d1 + j) * d2 + k) * d3;
        }
    }
}
template <class T> static void Deallocate(T *A, const int, const int = 0) { std::free(A); }
template <class T> static void Deallocate(T **A, const int, const int, const int = 0, const int = 0)
{
    if (!A) return;
    std::free(A[0]);
    std::free(A);
}
template <class T>
static void Deallocate(T ***A, const int, const int, const int, const int = 0, const int = 0, const int = 0)
{
    if (!A) return;
    std::free(A[0][0]);
    std::free(A[0]);
    std::free(A);
}
template <class T>
static void Deallocate(T ****A, const int, const int, const int, const int, const int = 0, const int = 0,
                       const int = 0, const int = 0)
{
    if (!A) return;
    std::free(A[0][0][0]);
    std::free(A[0][0]);
    std::free(A[0]);
    std::free(A);
}
