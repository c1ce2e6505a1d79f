// sg_fit.hip -- FitSKIRT's evaluation of a population of simulated frames against one reference
// image, on the device where the frames were detected:
//   ReferenceImage::chi2 (FitSKIRTcore/ReferenceImage.cpp:95-114), one-component case:
//   PSF convolution (Convolution::nested_loop, Convolution.cpp:34-66) with a normalised kernel
//   (GaussianKernel::setupSelfBefore, GaussianKernel.cpp:19-46; ConvolutionKernel::setupSelfAfter,
//   ConvolutionKernel.cpp:23-28), then the luminosity scale fitted by golden section
//   (GoldenSection::function and ::optimize, GoldenSection.cpp:57-124).
// One workgroup per member: the convolution writes the member's convolved frame, and the golden
// section loop evaluates both trial luminosities of an iteration in one pass over it.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/skirt7_amd.h"
#include "sg_host.hpp"

namespace sg {
namespace {

constexpr int kFitBlock = 256;

// Sum of (a, b) over the workgroup, returned to every thread; the order of the additions is fixed.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* s_part)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();   // the previous call's partial sums have been read
    if (__lane_id() == 0) { s_part[2 * w] = a; s_part[2 * w + 1] = b; }
    __syncthreads();
    a = 0; b = 0;
    for (int i = 0; i < kFitBlock / 64; i++) { a += s_part[2 * i]; b += s_part[2 * i + 1]; }
}

__global__ void __launch_bounds__(kFitBlock)
k_fit(const double* const* frames, double* work, const double* kern, int kx, int ky, const double* ref, int Nx,
      int Ny, double minlum, double maxlum, double* out)
{
    __shared__ double s_part[2 * (kFitBlock / 64)];
    const int mem = blockIdx.x;
    const int np = Nx * Ny;
    const double* in = frames[mem];
    double* cv = work + (size_t)mem * np;
    const int cx = (kx - 1) / 2, cy = (ky - 1) / 2;
    // Convolution::nested_loop scatters input pixel (xi, yi) times kernel(xk, yk) to output pixel
    // (xi - cx + xk, yi - cy + yk); gathered per output pixel, that is the same set of terms.
    // Pixels where the reference is zero are masked as GoldenSection::function does.
    for (int o = threadIdx.x; o < np; o += kFitBlock) {
        const int xo = o % Nx, yo = o / Nx;
        double acc = 0.0;
        for (int yk = 0; yk < ky; yk++) {
            const int yi = yo + cy - yk;
            if (yi < 0 || yi >= Ny) continue;
            for (int xk = 0; xk < kx; xk++) {
                const int xi = xo + cx - xk;
                if (xi < 0 || xi >= Nx) continue;
                acc += in[xi + Nx * yi] * kern[xk + kx * yk];
            }
        }
        cv[o] = ref[o] == 0.0 ? 0.0 : acc;
    }
    __syncthreads();
    // GoldenSection::optimize
    const double GOLD = 0.3819660113;
    double chi = 1.0, lumvalue = 1.0, prev = 1.0, a0 = minlum, b0 = maxlum;
    for (int j = 0; j < 300; j++) {
        const double d = (b0 - a0) * GOLD;
        const double a1 = a0 + d, b1 = b0 - d;
        double ca = 0.0, cb = 0.0;
        for (int o = threadIdx.x; o < np; o += kFitBlock) {
            const double r = ref[o];
            if (r == 0.0) continue;
            const double s = cv[o];
            const double ta = a1 * s, tb = b1 * s;
            const double qa = (r - ta) / sqrt(fabs(r) + ta);
            const double qb = (r - tb) / sqrt(fabs(r) + tb);
            ca += qa * qa;
            cb += qb * qb;
        }
        block_sum2(ca, cb, s_part);
        if (ca < cb) { b0 = b1; lumvalue = a1; chi = ca; }
        else { a0 = a1; lumvalue = b1; chi = cb; }
        if (fabs(prev - lumvalue) / lumvalue <= 1e-8 && j >= 20) break;
        prev = lumvalue;
    }
    if (threadIdx.x == 0) { out[2 * mem] = lumvalue; out[2 * mem + 1] = chi; }
}

} // namespace
} // namespace sg

using namespace sg;

extern "C" {

int sg_gaussian_kernel(double fwhm, int dimension, double* out)
{
    try {
        SG_REQUIRE(fwhm > 0, "FHWM should be positive");
        SG_REQUIRE(dimension > 0, "Frame dimension should be positive");
        const double sigma = fwhm / 2.3548;
        const int c = (dimension + 1) / 2;
        double sum = 0.0;
        for (int yk = 1; yk <= dimension; yk++)
            for (int xk = 1; xk <= dimension; xk++) {
                const int X = c - xk, Y = c - yk;
                double v = std::exp(-(X * X + Y * Y) / (2 * sigma * sigma));
                out[(xk - 1) + dimension * (yk - 1)] = v;
            }
        for (int i = 0; i < dimension * dimension; i++) sum += out[i];
        for (int i = 0; i < dimension * dimension; i++) out[i] /= sum;
        return 0;
    } catch (const Fail& f) {
        return f.code;
    }
}

int sg_fit_frames(int n, sg_instruments* const* instr, int instrument, int ell, const double* kernel, int kx,
                  int ky, const double* ref, double minlum, double maxlum, double* lum, double* chi2,
                  double* convolved)
{
    try {
        SG_REQUIRE(n >= 1 && instr, "a fit needs at least one member");
        SG_REQUIRE(kx >= 1 && ky >= 1 && kernel, "the convolution kernel needs at least one pixel");
        SG_REQUIRE(maxlum > minlum, "the luminosity interval is empty");
        SG_REQUIRE(instr[0] && instrument >= 0 && instrument < instr[0]->n, "instrument index out of range");
        const int Nx = instr[0]->dev[instrument].Nx, Ny = instr[0]->dev[instrument].Ny;
        const size_t np = (size_t)Nx * Ny;
        std::vector<const double*> fp(n);
        for (int i = 0; i < n; i++) {
            const sg_instruments* in = instr[i];
            SG_REQUIRE(in && instrument < in->n, "instrument index out of range");
            SG_REQUIRE(ell >= 0 && ell < in->nlambda, "wavelength index out of range");
            SG_REQUIRE(in->dev[instrument].Nx == Nx && in->dev[instrument].Ny == Ny,
                       "the members' frames must all have the reference image's size");
            fp[i] = in->frames[instrument]->as<double>() + (size_t)ell * np;
        }
        DevBuf dframes, dwork, dkern, dref, dout;
        dframes.upload(fp.data(), fp.size());
        dwork.alloc((size_t)n * np * sizeof(double));
        dkern.upload(kernel, (size_t)kx * ky);
        dref.upload(ref, np);
        dout.alloc((size_t)n * 2 * sizeof(double));
        hipLaunchKernelGGL(k_fit, dim3(n), dim3(kFitBlock), 0, 0, dframes.as<const double*>(), dwork.as<double>(),
                           dkern.as<double>(), kx, ky, dref.as<double>(), Nx, Ny, minlum, maxlum, dout.as<double>());
        SG_HIP(hipGetLastError());
        std::vector<double> res((size_t)n * 2);
        SG_HIP(hipMemcpy(res.data(), dout.p, res.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            if (lum) lum[i] = res[2 * i];
            if (chi2) chi2[i] = res[2 * i + 1];
        }
        if (convolved) SG_HIP(hipMemcpy(convolved, dwork.p, (size_t)n * np * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    } catch (const Fail& f) {
        return f.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -4;
    }
}

} // extern "C"
