// Rounding of gfx950's stochastic-rounding conversions on given inputs:
// out16[i] = v_cvt_sr_f16_f32(x[i], r[i]), outb[i] = v_cvt_sr_bf16_f32(x[i], r[i])
// (bit patterns).  Used by make_sr_vectors.py to record the vectors that pin
// the oracle's stochastic rounding.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef __bf16 b2_t __attribute__((ext_vector_type(2)));

__global__ void k_sr(const float* x, const uint32_t* r, uint16_t* o16, uint16_t* ob, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    h2_t z16 = {0, 0};
    b2_t zb = {0, 0};
    o16[i] = __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_f16_f32(z16, x[i], r[i], false)[0]);
    ob[i] = __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_bf16_f32(zb, x[i], r[i], false)[0]);
}

extern "C" int sr_probe(const float* x, const uint32_t* r, uint16_t* o16, uint16_t* ob, int n) {
    if (n <= 0) return 1;
    float* dx = nullptr;
    uint32_t* dr = nullptr;
    uint16_t *d16 = nullptr, *db = nullptr;
    int rc = 0;
    if (hipMalloc(&dx, (size_t)n * 4) != hipSuccess || hipMalloc(&dr, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&d16, (size_t)n * 2) != hipSuccess || hipMalloc(&db, (size_t)n * 2) != hipSuccess)
        rc = 2;
    if (!rc && (hipMemcpy(dx, x, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dr, r, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess))
        rc = 3;
    if (!rc) {
        k_sr<<<(n + 255) / 256, 256>>>(dx, dr, d16, db, n);
        if (hipDeviceSynchronize() != hipSuccess) rc = 4;
    }
    if (!rc && (hipMemcpy(o16, d16, (size_t)n * 2, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(ob, db, (size_t)n * 2, hipMemcpyDeviceToHost) != hipSuccess))
        rc = 5;
    if (dx) (void)hipFree(dx);
    if (dr) (void)hipFree(dr);
    if (d16) (void)hipFree(d16);
    if (db) (void)hipFree(db);
    return rc;
}
