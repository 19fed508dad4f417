/*
 * amr_internal.h -- helpers shared by the library's translation units; not
 * part of the C ABI (hidden visibility).
 */
#ifndef AMR_INTERNAL_H
#define AMR_INTERNAL_H

#include <stddef.h>

#include "amr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AMR_HIDDEN __attribute__((visibility("hidden")))

struct amrgpu_ctx;
struct amr_forest;

/* Formats the message returned by amrgpu_last_error(). */
AMR_HIDDEN void amr_set_error(const char *fmt, ...);

/* Device context for the whole forest with the current physics settings
 * (one instance; h_base = coarse_grid_h, walls bc[0..3]), or with the given
 * instance settings when override is not NULL. */
AMR_HIDDEN int amr_device_setup(struct amr_forest *f, const amrgpu_inst_t *override, struct amrgpu_ctx **out);
/* Regrid the forest by tags (amr_regrid_tags) and remap the device state;
 * wall seconds of the host forest part and of the device remesh are added to
 * *t_forest and *t_remesh (either may be NULL). */
AMR_HIDDEN int amr_device_regrid(struct amr_forest *f, struct amrgpu_ctx *g, const int *tags,
                                 int allow_coarsening, double *t_forest, double *t_remesh);
AMR_HIDDEN double amr_now(void);
/* The context's staging buffer (device, *ndoubles doubles: one state vector
 * of the current capacity), free between uploads and downloads; kernels on
 * the context's stream may use it as scratch. */
AMR_HIDDEN int amrgpu_staging(struct amrgpu_ctx *c, double **buf, size_t *ndoubles);
/* Per-quadrant loops of the host forest run on OpenMP threads unless turned
 * off for the calling thread (the batch's per-instance regrids do).  Returns
 * the thread's previous setting, which the caller restores. */
AMR_HIDDEN int amr_forest_host_parallel(int on);
/* The calling thread's current HIP device, and making a device current on
 * the calling thread (threads the library starts use the caller's device). */
AMR_HIDDEN int amrgpu_current_device(void);
AMR_HIDDEN void amrgpu_use_device(int dev);

#ifdef __cplusplus
}
#endif
#endif
