// atomics.h -- non-returning float/double atomic adds with an explicit
// address space, so the compiler emits ds_add_f32/f64 (LDS) and
// global_atomic_add_f32/f64 (HBM) instead of generic flat atomics, which it
// falls back to whenever a pointer's address space is not provable (pointer
// arrays, lambda captures, offsets into extern __shared__).
#pragma once
#include <hip/hip_runtime.h>

namespace pic {

typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(1))) float glb_f32;
typedef __attribute__((address_space(1))) double glb_f64;

__device__ __forceinline__ void lds_add(float* p, float v) {
    __hip_atomic_fetch_add((lds_f32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add((lds_f64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
__device__ __forceinline__ void lds_add_u64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add((lds_u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
typedef __attribute__((address_space(3))) int lds_i32;
__device__ __forceinline__ void lds_add_i32(int* p, int v) {
    __hip_atomic_fetch_add((lds_i32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void hbm_add(float* p, float v) {
    __hip_atomic_fetch_add((glb_f32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void hbm_add(double* p, double v) {
    __hip_atomic_fetch_add((glb_f64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace pic
