// flr_waveform.hip -- galactic-binary waveforms generated on the device, straight into a
// packed response input (complex128, complex64 h / s, or 32-bit fixed point h / s), so that
// batches and served requests can be described by source parameters instead of waveform
// arrays. The formula is the reference test's generator (tests/test_fastlisaresponse.py:41-63):
//   phase = 2 pi (f t + fdot t^2 / 2 + fddot t^3 / 6) - phi0
//   h+_s = -cos(phase) A (1 + cos^2 i),  hx_s = -sin(phase) 2 A cos i,
//   h+ = h+_s cos 2psi - hx_s sin 2psi,  hx = h+_s sin 2psi + hx_s cos 2psi,
// with t = t0 + k dt, evaluated in FP64 in that operation order (the library is built with
// -ffp-contract=off; the rounding of 2 pi times a phase of many cycles is the reference's too,
// and the null channel T is sensitive to it; sin and cos of the phase come from phase_sincos,
// within ~2e-16 of the exact values). cos 2psi, sin 2psi and cos i come from the host.
#include "flr_internal.hpp"
#include "flr_sincos.hpp"

namespace flr {

// Source of packed sample g: the last source whose offset is <= g (sources sorted by offset).
__device__ __forceinline__ int gb_source_of(const flr_gb_source *s, int n, int64_t g)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s[mid].offset <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

constexpr int kGbPerThread = 4;  // samples per thread: one source lookup per 1024 samples

// Sample k of source s into packed slot g of `out`.
__device__ __forceinline__ void gb_write(const flr_gb_source &s, int64_t k, int64_t g, double dt,
                                         int format, void *__restrict__ out)
{
    const double t = s.t0 + (double)k * dt;
    const double two_pi = 6.283185307179586;  // 2 * pi in FP64
    const double tt = t * t;
    const double phase =
        two_pi * ((s.f0 * t + 0.5 * s.fdot * tt) + (1.0 / 6.0) * s.fddot * (tt * t)) - s.phi0;
    double sn, cs;
    phase_sincos(phase, &sn, &cs);
    const double hSp = -cs * s.amplitude * (1.0 + s.cos_iota * s.cos_iota);
    const double hSc = -sn * 2.0 * s.amplitude * s.cos_iota;
    const double hp = hSp * s.cos2psi - hSc * s.sin2psi;
    const double hc = hSp * s.sin2psi + hSc * s.cos2psi;
    if (format == FLR_GB_COMPLEX128) {
        reinterpret_cast<double2 *>(out)[g] = make_double2(hp, hc);
    } else if (format == FLR_GB_COMPLEX64) {
        reinterpret_cast<float2 *>(out)[g] = make_float2((float)(hp / s.scale), (float)(hc / s.scale));
    } else {
        const double full = 2147483647.0;  // 2^31 - 1
        const double qr = fmin(fmax(rint(hp / s.scale * full), -full), full);
        const double qi = fmin(fmax(rint(hc / s.scale * full), -full), full);
        reinterpret_cast<int2 *>(out)[g] = make_int2((int)qr, (int)qi);
    }
}

__global__ __launch_bounds__(kBlock) void gb_kernel(const flr_gb_source *__restrict__ src,
                                                    int num_sources, int64_t total, double dt,
                                                    int format, void *__restrict__ out)
{
    // the workgroup's first sample's source, found once
    __shared__ int first_source;
    const int64_t g0 = (int64_t)blockIdx.x * (kBlock * kGbPerThread);
    if (threadIdx.x == 0) first_source = gb_source_of(src, num_sources, g0);
    __syncthreads();
    const int s0 = __builtin_amdgcn_readfirstlane(first_source);
    {
        // the usual case, the workgroup's samples inside that one source: its record is read
        // once per wave (a uniform address)
        const flr_gb_source S = src[s0];
        const int64_t g1 = g0 + kBlock * kGbPerThread;
        if (g0 >= S.offset && g1 <= S.offset + S.num_pts && g1 <= total) {
#pragma unroll 1
            for (int r = 0; r < kGbPerThread; ++r) {
                const int64_t g = g0 + (int64_t)r * kBlock + threadIdx.x;
                gb_write(S, g - S.offset, g, dt, format, out);
            }
            return;
        }
    }
    // otherwise each thread steps forward from that source (a 1024-sample range spans few)
    int si = s0;
#pragma unroll 1
    for (int r = 0; r < kGbPerThread; ++r) {
        const int64_t g = g0 + (int64_t)r * kBlock + threadIdx.x;
        if (g >= total) return;
        while (si + 1 < num_sources && src[si + 1].offset <= g) ++si;
        const flr_gb_source s = src[si];
        const int64_t k = g - s.offset;
        if (k < 0 || k >= s.num_pts) continue;  // a gap between sources: left untouched
        gb_write(s, k, g, dt, format, out);
    }
}

hipError_t launch_gb(const flr_gb_source *sources, int num_sources, int64_t total, double dt,
                     int format, void *out, hipStream_t stream)
{
    if (total <= 0) return hipSuccess;
    const int64_t per_block = (int64_t)kBlock * kGbPerThread;
    const int64_t blocks = (total + per_block - 1) / per_block;
    hipLaunchKernelGGL(gb_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, stream, sources,
                       num_sources, total, dt, format, out);
    return hipGetLastError();
}

}  // namespace flr
