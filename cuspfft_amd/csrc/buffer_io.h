// buffer_io.h -- raw buffer loads and stores for the row-pass kernels.
//
// A buffer descriptor (base, size in bytes) lets an access split its offset
// into a per-lane VGPR part and a wave-uniform scalar part (no 64-bit address
// per access), and the hardware bounds check -- on gfx950 it tests the sum of
// both parts against the size -- makes reads past the end return 0 and drops
// writes past the end.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spfft {

namespace bio {

using rsrc_t = __amdgpu_buffer_rsrc_t;
// raw buffer: stride 0, num_records in bytes; word 3 = 32-bit data format (gfx9)
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kRsrcWord3);
}
// An offset at or past num_records makes a load return 0 and drops a store;
// kOob is past every buffer made here (num_records <= 2^31 - 256).
constexpr int kOob = 0x7fffff00;

// The LLVM buffer intrinsics, bound by name: this compiler lowers the clang
// builtins __builtin_amdgcn_raw_buffer_load_b64 / _b128 to a single 32-bit load.
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ f32x2 raw_load_v2f32(rsrc_t r, int voff, int soff, int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ void raw_store_v2f32(f32x2 v, rsrc_t r, int voff, int soff, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.v2f32");

// Cache policy of the accesses made once per transform (the pass-1 rows, the
// tile buffer, the output): non-temporal (bit 1 of the aux operand on gfx950),
// so they do not evict the tables every row reuses from L2.
constexpr int kStream = 2;

__device__ __forceinline__ float2 load2(rsrc_t r, int voff, int soff) {
    const f32x2 d = raw_load_v2f32(r, voff, soff, 0);
    return make_float2(d[0], d[1]);
}
__device__ __forceinline__ float2 load2_stream(rsrc_t r, int voff, int soff) {
    const f32x2 d = raw_load_v2f32(r, voff, soff, kStream);
    return make_float2(d[0], d[1]);
}
__device__ __forceinline__ void store2_stream(float2 x, rsrc_t r, int voff, int soff) {
    const f32x2 d = {x.x, x.y};
    raw_store_v2f32(d, r, voff, soff, kStream);
}

}  // namespace bio

}  // namespace spfft
