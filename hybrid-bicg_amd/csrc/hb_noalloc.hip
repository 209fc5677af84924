/*
 * hb_noalloc.hip -- hbicg_plan_true_residuals_noalloc: the plan's true
 * residual check without any device allocation, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops).  It has no
 * counterpart in the reference.
 *
 * hbicg_plan_true_residuals allocates its buffer of workgroup partial sums
 * on every call.  This entry point runs the same kernels with the partial
 * sums in `scratch`, a static array in device memory: every GPU the library
 * runs on holds its own 1 MiB copy until the process ends.  The kernels get
 * its device address, looked up once per device with hipGetSymbolAddress
 * (the array's name in host code is not a device address).  Every partial
 * pair the combining kernel reads is written first by the same call, so no
 * contents of an earlier call are ever read; calls that overlap on one GPU
 * share the array and race.
 */
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdio.h>

#include "hb_launch.h"
#include "hb_plan.h"
#include "hb_shared.h"

/* 2 * HB_RESID_PAIRS doubles: the partial pairs of one evaluation of up to
 * HB_RESID_PAIRS shifts (hb_k_resid_blocks keeps M * R within it). */
static __device__ __align__(256) double scratch[2 * HB_RESID_PAIRS];

#define HB_MAX_DEVICES 64
static std::once_flag g_once[HB_MAX_DEVICES];
static double *g_ptr[HB_MAX_DEVICES];
static hipError_t g_err[HB_MAX_DEVICES];

/* This device's copy of scratch; dev must be the current device. */
static double *scratch_ptr(int dev, hipError_t *err)
{
    if (dev < 0 || dev >= HB_MAX_DEVICES) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::call_once(g_once[dev], [dev]() {
        void *p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = g_err[dev] == hipSuccess ? (double *)p : nullptr;
    });
    *err = g_err[dev];
    return g_ptr[dev];
}

extern "C" int hbicg_plan_true_residuals_noalloc(hbicg_plan *pl, const double *b, const double *sigma,
                                                 const double *x_set, double *out)
{
    if (!pl || !b || !sigma || !x_set || !out) return HBICG_E_ARG;
    const hb_dev *d = hb_plan_dev(pl);
    if (d->M > HB_RESID_PAIRS) {
        fprintf(stderr, "hbicg_plan_true_residuals_noalloc: %d shifts exceed the %d its scratch holds\n",
                d->M, HB_RESID_PAIRS);
        return HBICG_E_SCRATCH;
    }
    const int dev = hb_plan_device(pl);
    if (hipSetDevice(dev) != hipSuccess) return HBICG_E_HIP;
    hipError_t err;
    double *part = scratch_ptr(dev, &err);
    if (!part) {
        fprintf(stderr, "hbicg_plan_true_residuals_noalloc: hipGetSymbolAddress: %s\n", hipGetErrorString(err));
        return HBICG_E_HIP;
    }
    if (hb_k_true_residual(d, b, sigma, x_set, part, out, hb_plan_stream(pl)) != 0) return HBICG_E_HIP;
    return 0;
}
