/*
 * HiG-Flow step on an MI355X through libhigflow_hip: the driver-side binding.
 *
 * A driver attaches a set-up higflow_solver once; from then on
 * higflow_solver_step / higflow_solver_step_viscoelastic run on the GPU
 * through a one-line branch at the top of each:
 *     if (higflow_hip_step(ns) == 0) return;
 */
#ifndef HIGFLOW_STEP_HIP_H
#define HIGFLOW_STEP_HIP_H

#include "hig-flow-kernel.h"

/* Mirror `ns` on the GPU.  Call after the driver's higflow_initialize_*
 * and higflow_create_partitioned_* calls, on one rank.  The domain and
 * boundary .amr files are the ones the driver's <prefix>.domain.yaml and
 * <prefix>.bc.yaml name, boundaries in file order with their user ids.
 * Returns 0, or -1 when the configuration has no GPU path (the driver keeps
 * the CPU path; hfh_last_error() says why). */
int higflow_hip_attach(higflow_solver *ns, int ndomains, const char *const domain_amr[], int nbcs,
                       const int bc_id[], const char *const bc_amr[]);
void higflow_hip_detach(higflow_solver *ns);
/* dp->values -> device (after the driver sets or loads fields) and back
 * (before host IO: higflow_print_vtk, higflow_save_properties, probes) */
int higflow_hip_push(higflow_solver *ns);
int higflow_hip_pull(higflow_solver *ns);
/* One step on the GPU when `ns` is attached: returns 0, or -1 when it is
 * not (the caller runs its CPU step). */
int higflow_hip_step(higflow_solver *ns);

#endif
