// tdc_serve.h -- device-side tables of the serving form (shared host/device).
#pragma once

#include <hip/hip_runtime.h>

#include "tdc_internal.h"

namespace tdc {

// Pixels of one serving tile: one per thread of the workgroup (the matrix
// engine's tile for the precision form, one pixel per lane for the plain form's
// per-chirp range tile).
constexpr int kServeTilePixels = kBlock;
// Capacity of one step: tiles, finalize blocks, and runs of one job's tiles
// (a workgroup reads a step's runs of its tile family into LDS once).
constexpr int kMaxStepTiles = 32768;
constexpr int kMaxStepFin = 32768;
constexpr int kMaxStepRuns = 256;
constexpr int kMaxJobSlots = 8192;

// One admitted job as the kernels see it.
struct alignas(16) ServeJob {
    const void *raw;          // precision: range-profile table or matrix engine A tiles;
                              // plain: (rr, ri, dk_m, 0) terms
    const float *inv_scale;   // precision: per-chirp scale of the A tiles
    const ChirpPos *cpos;     // FP64 phase centre per chirp (range profile: from the origin
                              // pp.o, with its squared norm in pad)
    const double *grid64;     // FP64 pixel coordinates
    const double *kanchor;    // matrix engine: 2 (f0 + K t_a) / c per 64-sample block;
                              // plain: the anchor wavenumbers k_b per 64-sample block
    void *partial;            // [nsplit][npx]: double2 (matrix engine) or float2 (plain form,
                              // range-profile engine)
    float2 *image;            // [npx]
    double dk;                // precision, matrix engine: 2 K dt / c
    ProfileParams pp;         // precision, range-profile engine
    int npx, npri, ns, nsplit, cps, form, storage, profile;
    int npt;                  // pixel tiles of kServeTilePixels pixels
    int pad;
};

// A run of consecutive tiles of one job in a step: the step's tiles
// [first, first + count) of the run's family are the job's own tiles
// [tile0, tile0 + count), job tile t being pixel tile t % npt of chirp range
// t / npt. A run names its job by the device address of the job's
// descriptor, which travels to HBM in the job's own upload.
struct StepRun {
    const ServeJob *job;
    int first;
    int count;
};

struct FinTile {
    const ServeJob *job;
    int pix0; // first of up to kBlock pixels
    int pad;
};

struct StepTable {
    int ntiles[2]; // tiles of family 0 (plain and range-profile) and 1 (matrix engine)
    int nruns[2];  // runs[0, nruns[0]): family 0; [nruns[0], nruns[0] + nruns[1]): family 1
    int nfin;
    int pad;
    // the values the two tile counters hold when this step starts: counters
    // are never reset, each step takes tile t at counter value base + t
    unsigned int base[2];
    StepRun runs[kMaxStepRuns];
    int tile0[kMaxStepRuns]; // the job tile of each run's first tile
    FinTile fin[kMaxStepFin];
};

// `counter`: two counters, one per tile family, read relative to the step
// table's bases. `matrix`: also launch the matrix-engine family's kernel (a
// step with no such tiles leaves it out). Every workgroup of a family's
// kernel advances its counter by one beyond the tiles it takes.
hipError_t launch_serve_step(const StepTable *table, unsigned int *counter, int workgroups,
                             bool matrix, hipStream_t s);
hipError_t launch_serve_finalize(const StepTable *table, int workgroups, hipStream_t s);
// Plain-form terms (rr, ri, dk_m, 0) of a served job from its complex64 raw
// samples (windows applied) and the per-sample dk_m.
hipError_t launch_stage_plain_chirp(const float *raw, const float *dk, int npri, int ns,
                                    float4 *terms, hipStream_t s);

} // namespace tdc
