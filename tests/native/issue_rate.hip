// issue_rate: measures the VALU issue ceilings that bound the pair kernels on
// this GPU, with the same compiler flags as libncg (packed FP32 off).
//
//   fma32 / add32 / mul32 / fma64 / rsq32: independent chains (ILP 8-16) of
//     one instruction kind, 16 workgroups of 256 threads per CU;
//   far: the FP32 far-field interaction of the mixed pair kernel (3 adds,
//     3 for r^2, v_rsq_f32, 3 multiplies, 3 FMAs), j records read with
//     wave-uniform addresses from a 128-entry tile as in the kernel.
//   add64 / add64+cvt+add32: FP64 adds alone, and each followed by its
//     rounding to FP32 and an FP32 add (the separations of the ensemble's
//     mixed form): the difference prices v_cvt_f32_f64.
//
//   chain64 / chain32: one wavefront, one dependent chain of adds (the
//     body-order sums of the exact mesh field are such chains): the
//     latency of a dependent v_add_f64 / v_add_f32.
//
// Output: wave-instructions per CU per clock (at the reported clock) and,
// for `far`, interactions per second: the ceiling of the far-field loop;
// for the chains, clocks per dependent add.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s: %s (line %d)\n", #x, hipGetErrorString(e_), __LINE__); \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

namespace {

constexpr int WG = 256;

template <int ILP>
__global__ __launch_bounds__(WG) void k_fma32(float *out, int iters, float b, float c)
{
    float a[ILP];
    for (int q = 0; q < ILP; ++q) a[q] = threadIdx.x * 1e-3f + q;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < ILP; ++q) a[q] = fmaf(a[q], b, c);
    }
    float s = 0.f;
    for (int q = 0; q < ILP; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

__global__ __launch_bounds__(WG) void k_add32(float *out, int iters, float b)
{
    float a[8];
    for (int q = 0; q < 8; ++q) a[q] = threadIdx.x * 1e-3f + q;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = a[q] + b;
    }
    float s = 0.f;
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

__global__ __launch_bounds__(WG) void k_mul32(float *out, int iters, float b)
{
    float a[8];
    for (int q = 0; q < 8; ++q) a[q] = threadIdx.x * 1e-3f + q;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = a[q] * b;
    }
    float s = 0.f;
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

__global__ __launch_bounds__(WG) void k_fma64(double *out, int iters, double b, double c)
{
    double a[8];
    for (int q = 0; q < 8; ++q) a[q] = threadIdx.x * 1e-3 + q;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = fma(a[q], b, c);
    }
    double s = 0.;
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

// 8 chains of (v_rsq_f32, v_add_f32).
__global__ __launch_bounds__(WG) void k_rsq32(float *out, int iters, float b)
{
    float a[8];
    for (int q = 0; q < 8; ++q) a[q] = threadIdx.x * 1e-3f + q + 1.f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = __builtin_amdgcn_rsqf(a[q]) + b;
    }
    float s = 0.f;
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

__global__ __launch_bounds__(WG) void k_far(const float4 *__restrict__ ju, int iters, float *out)
{
    float wx = threadIdx.x * 1e-4f + 0.3f, wy = 0.2f, wz = 0.1f;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll 8
        for (int jj = 0; jj < 128; ++jj) {
            const float4 u = ju[jj];
            const float dx = wx + u.x, dy = wy + u.y, dz = wz + u.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const float y = __builtin_amdgcn_rsqf(r2);
            const float s3 = (u.w * y) * (y * y);
            fx = fmaf(s3, dx, fx);
            fy = fmaf(s3, dy, fy);
            fz = fmaf(s3, dz, fz);
        }
        wx += 1e-7f;
    }
    out[blockIdx.x * WG + threadIdx.x] = fx + fy + fz;
}

__global__ __launch_bounds__(WG) void k_add64(double *out, int iters, double b)
{
    double a[8];
    for (int q = 0; q < 8; ++q) a[q] = threadIdx.x * 1e-3 + q;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = a[q] + b;
    }
    double s = 0.;
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

// 8 chains of (v_add_f64, v_cvt_f32_f64, v_add_f32): the FP64 separation
// rounded to FP32 of the ensemble's mixed form.
__global__ __launch_bounds__(WG) void k_cvt(float *out, int iters, double b)
{
    double d[8];
    float a[8];
    for (int q = 0; q < 8; ++q) {
        d[q] = threadIdx.x * 1e-3 + q;
        a[q] = 0.f;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            d[q] = d[q] + b;
            a[q] += (float)d[q];
        }
    }
    float s = 0.f;
    for (int q = 0; q < 8; ++q) s += a[q];
    out[blockIdx.x * WG + threadIdx.x] = s;
}

template <class T>
__global__ __launch_bounds__(64) void k_chain(T *out, int iters, T b)
{
    T a = (T)threadIdx.x;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 32; ++q) a = a + b;
    }
    out[threadIdx.x] = a;
}

}  // namespace

int main()
{
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    const double ghz = p.clockRate * 1e-6;
    printf("device %s, %d CUs, clock %.3f GHz\n", p.gcnArchName, cus, ghz);
    const int blocks = cus * 16;
    float *o32 = nullptr, *hj = nullptr;
    double *o64 = nullptr;
    float4 *ju = nullptr;
    CK(hipMalloc(&o32, sizeof(float) * blocks * WG));
    CK(hipMalloc(&o64, sizeof(double) * blocks * WG));
    CK(hipMalloc(&ju, sizeof(float4) * 128));
    std::vector<float> h(512);
    for (int k = 0; k < 128; ++k) {
        h[4 * k] = 0.01f * k;
        h[4 * k + 1] = -0.003f * k;
        h[4 * k + 2] = 0.002f;
        h[4 * k + 3] = 1e-9f;
    }
    CK(hipMemcpy(ju, h.data(), sizeof(float) * 512, hipMemcpyHostToDevice));
    (void)hj;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int iters = 4096, far_iters = 64;
    const double waves = blocks * (WG / 64.0);
    // kind: 0 fma32 x16, 1 add32, 2 mul32, 3 fma64, 4 rsq32, 5 far, 6 add64,
    // 7 add64+cvt+add32
    const char *names[8] = {"fma32", "add32", "mul32", "fma64", "rsq32+add32", "far", "add64",
                            "add64+cvt+add32"};
    const double instr[8] = {16.0 * iters, 8.0 * iters, 8.0 * iters, 8.0 * iters, 16.0 * iters,
                             13.0 * 128 * far_iters, 8.0 * iters, 24.0 * iters};
    for (int kind = 0; kind < 8; ++kind) {
        auto launch = [&]() {
            switch (kind) {
            case 0: hipLaunchKernelGGL(k_fma32<16>, dim3(blocks), dim3(WG), 0, 0, o32, iters, 0.999f, 1e-3f); break;
            case 1: hipLaunchKernelGGL(k_add32, dim3(blocks), dim3(WG), 0, 0, o32, iters, 1e-3f); break;
            case 2: hipLaunchKernelGGL(k_mul32, dim3(blocks), dim3(WG), 0, 0, o32, iters, 0.999f); break;
            case 3: hipLaunchKernelGGL(k_fma64, dim3(blocks), dim3(WG), 0, 0, o64, iters, 0.999, 1e-3); break;
            case 4: hipLaunchKernelGGL(k_rsq32, dim3(blocks), dim3(WG), 0, 0, o32, iters, 0.5f); break;
            case 6: hipLaunchKernelGGL(k_add64, dim3(blocks), dim3(WG), 0, 0, o64, iters, 1e-3); break;
            case 7: hipLaunchKernelGGL(k_cvt, dim3(blocks), dim3(WG), 0, 0, o32, iters, 1e-3); break;
            default: hipLaunchKernelGGL(k_far, dim3(blocks), dim3(WG), 0, 0, ju, far_iters, o32); break;
            }
        };
        launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < 5; ++r) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        ms /= 5.f;
        const double rate = waves * instr[kind] / (ms * 1e-3);
        printf("%-12s %8.3f ms  %.3e wave-instr/s  %.3f per CU per clock", names[kind], ms, rate,
               rate / cus / (ghz * 1e9));
        if (kind == 5) {
            const double inter = (double)blocks * WG * 128 * far_iters / (ms * 1e-3);
            printf("  %.3e interactions/s (%.1f TFLOP/s at 20 flop)", inter, 20.0 * inter / 1e12);
        }
        printf("\n");
    }
    for (int kind = 0; kind < 2; ++kind) {
        const int citers = 1 << 15;
        auto launch = [&]() {
            if (kind == 0) hipLaunchKernelGGL(k_chain<double>, dim3(1), dim3(64), 0, 0, o64, citers, 1e-3);
            else hipLaunchKernelGGL(k_chain<float>, dim3(1), dim3(64), 0, 0, o32, citers, 1e-3f);
        };
        launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double adds = 32.0 * citers;
        printf("%-12s %8.3f ms  %.3f ns per dependent add  %.2f clocks\n", kind ? "chain32" : "chain64",
               ms, ms * 1e6 / adds, ms * 1e-3 / adds * ghz * 1e9);
    }
    CK(hipFree(o32));
    CK(hipFree(o64));
    CK(hipFree(ju));
    return 0;
}
