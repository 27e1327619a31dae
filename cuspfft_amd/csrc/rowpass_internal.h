// rowpass_internal.h -- shared pieces of the row-pass translation units.
#pragma once
#include "codelets.h"
#include "spfft_device.h"
#include "spfft_kernels.h"

namespace spfft {

constexpr int kThreads = 256;

// Lengths M1 with instantiated k_fft1 / k_ifft1, in three groups (one per
// translation unit, so they compile in parallel).
#define SPFFT_M1_GROUP_A(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12) X(14) X(15) X(16)
#define SPFFT_M1_GROUP_B(X) X(18) X(20) X(21) X(24) X(25) X(27) X(28) X(30) X(32)
#define SPFFT_M1_GROUP_C(X) X(35) X(36) X(40) X(42) X(45) X(48) X(49) X(54) X(56) X(60) X(63) X(64)
#define SPFFT_M1_LIST(X) SPFFT_M1_GROUP_A(X) SPFFT_M1_GROUP_B(X) SPFFT_M1_GROUP_C(X)

// Butterflies a thread may hold across the load/store barrier of one k_conv stage
// (at most 32 complex values in registers).
template <int R> struct MaxBfly { static constexpr int value = R >= 32 ? 1 : 32 / R; };

// Each returns hipErrorInvalidValue when M1 is not in its group.
hipError_t launch_fft1_group_a(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse);
hipError_t launch_fft1_group_b(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse);
hipError_t launch_fft1_group_c(const RowPassArgs& a, dim3 grid, hipStream_t s, bool inverse);
// k_conv launchers, one per translation unit: each returns true (and sets
// *err) when the plan's stage radices belong to it.
bool mfma_supports_m1(int m1);
hipError_t launch_fft1_mfma(const RowPassArgs& a, hipStream_t s, bool inverse);
bool launch_conv_a(const RowPassArgs& a, hipStream_t s, hipError_t* err);
bool launch_conv_b(const RowPassArgs& a, hipStream_t s, hipError_t* err);
bool launch_conv_c(const RowPassArgs& a, hipStream_t s, hipError_t* err);
bool launch_conv_d(const RowPassArgs& a, hipStream_t s, hipError_t* err);
bool launch_conv_e(const RowPassArgs& a, hipStream_t s, hipError_t* err);

}  // namespace spfft
