// vox.h -- per-voxel access to the component-major (SoA) fields through buffer
// resources.
//
// A field component column is a[c*ld + v] (ld: the column stride).  Addressed as a plain pointer, every
// column of every field costs a 64-bit per-lane address, and the compiler keeps
// (or spills) one address pair per column it loads.  A buffer resource is
// wave-uniform (scalar registers), so each lane needs only its 32-bit byte offset
// v*8, shared by every column of every field.  One resource covers a group of 8
// columns (the column within the group goes into the scalar offset), which keeps
// the number of live 4-register resources, and their spills, down; its record
// count is the group's length, so an access past the group reads 0 and a store
// there is dropped.  Needs 64 ld < 2^32 (ld < 2^26).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cg {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

struct Vox {
    uint32_t off;  // v * 8
    uint32_t nb;   // ld * 8 (bytes of one column)
    long cs;       // column stride ld (doubles)
    static constexpr int GC = 8;  // columns per resource

    __device__ __forceinline__ Vox(long v, long ld_) : off((uint32_t)v * 8u), nb((uint32_t)(ld_ * 8)), cs(ld_) {}

    __device__ __forceinline__ static __amdgpu_buffer_rsrc_t rs(const void* p, uint32_t nbytes) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000);
    }
    // component c of field a at this voxel
    __device__ __forceinline__ double ld(const double* a, long c) const {
        const long g = c & ~(long)(GC - 1);
        const uint32_t so = (uint32_t)(c - g) * nb;
        return __builtin_bit_cast(double,
                                  __builtin_amdgcn_raw_buffer_load_b64(rs(a + g * cs, nb * GC), (int)off, (int)so, 0));
    }
    __device__ __forceinline__ void st(double* a, long c, double x) const {
        const long g = c & ~(long)(GC - 1);
        const uint32_t so = (uint32_t)(c - g) * nb;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), rs(a + g * cs, nb * GC), (int)off, (int)so,
                                              0);
    }
};

}  // namespace cg
