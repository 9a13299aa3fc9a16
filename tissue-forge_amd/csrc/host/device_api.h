// Interface between the host engine and the device module (csrc/device).
// Everything here is host-callable C++; the kernels stay private to the
// device module.
#pragma once

#include <cstdint>

namespace tfg {

class Engine;
struct Energies;
struct DeviceEnsemble;

enum Precision : int { PREC_MIXED = 0, PREC_FP64 = 1 };

int device_select(int ordinal);
int device_current();   // the calling thread's device, -1 without one
int device_count();

// Uploads the engines as independent systems of one ensemble.  Every engine
// keeps a pointer to the ensemble until ensemble_destroy.
int ensemble_create(DeviceEnsemble **out, Engine *const *engines, int n, int precision);
int ensemble_destroy(DeviceEnsemble *ens, bool download);
// Copies the device state of every system back into its host engine.
int ensemble_download(DeviceEnsemble *ens);
int ensemble_download_one(DeviceEnsemble *ens, int sys);
// Whole steps: binning, nonbonded, bonded, flux, forward Euler.
int ensemble_step(DeviceEnsemble *ens, int nsteps, bool use_graph);
// Binning + nonbonded + bonded force evaluation only (no update).
int ensemble_force(DeviceEnsemble *ens);
int ensemble_energies(DeviceEnsemble *ens, int sys, Energies *out);
// engine_force for a host engine in one call that allocates no device
// memory: every buffer comes from the current GPU's fixed scratch array
// (scratch_bytes() bytes).  Forces and number density by id into f (n*4).
int nonbond_eval_no_alloc(Engine *e, int precision, double *f, Energies *en);
size_t scratch_bytes();
// FP64 per-species totals of one system after the last step; returns count.
int ensemble_species_totals(DeviceEnsemble *ens, int sys, double *out, int cap);
int ensemble_set_launch(DeviceEnsemble *ens, int blocks, int threads);
int ensemble_get_launch(DeviceEnsemble *ens, int *blocks, int *threads);
// Times candidate grids of the pair kernel and of the streaming kernels on
// force-only passes and keeps the fastest (results do not depend on them).
int ensemble_autotune(DeviceEnsemble *ens, int reps, int *blocks, int *stream_blocks);
// Per-kernel timing of the last ensemble_step / ensemble_force call.
int ensemble_timing(DeviceEnsemble *ens, double *ms_out, int cap, const char **names_out);
// Milliseconds of device time for the last nsteps of ensemble_step.
double ensemble_last_ms(DeviceEnsemble *ens);
int ensemble_precision(DeviceEnsemble *ens);
int ensemble_size(DeviceEnsemble *ens);
// Per-step phase timing with HIP events on the module's stream (eager steps).
int ensemble_set_profile(DeviceEnsemble *ens, bool on);
// Neighbour lists: in use, rebuilds of system sys so far, entries per particle.
int ensemble_list_state(DeviceEnsemble *ens, int sys, int *active, long long *rebuilds, int *capacity);
// Mean and largest list length of system sys at its last build.
int ensemble_list_entries(DeviceEnsemble *ens, int sys, double *mean, int *maxn);

}  // namespace tfg
