// The phased engine: the reference's per-phase kernels (Collision, pack /
// exchange / unpack, Streaming with its boundary kernels, Update) on the
// reference's ghost-3 layout, FP64.  It carries the reference's own boundary
// code, quirks included, for all three cases; the fused engine carries the
// fast path.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/imexlb.h"

namespace imx {

struct PhasedState;

int phased_create(const imexlb_config &c, hipStream_t st, PhasedState **out, std::string &err);
void phased_destroy(PhasedState *s);
int64_t phased_device_bytes(const PhasedState *s);
int phased_initialize(PhasedState *s, std::string &err);
int phased_collide(PhasedState *s, std::string &err);
int phased_pack(PhasedState *s, std::string &err);
int phased_exchange(PhasedState *s, std::string &err);
int phased_unpack(PhasedState *s, std::string &err);
int phased_stream(PhasedState *s, std::string &err);
int phased_update(PhasedState *s, std::string &err);
int phased_full_step(PhasedState *s, std::string &err);
// n full steps, fused where the phases allow; same results as n full steps
int phased_steps(PhasedState *s, int n, std::string &err);
int phased_get_moments(PhasedState *s, double *p, double *u, double *v, double *w, std::string &err);
int phased_get_f(PhasedState *s, double *f, std::string &err);
int phased_set_f(PhasedState *s, const double *f, std::string &err);
int phased_diagnostics(PhasedState *s, double *ke, double *mass, std::string &err);
// whole state for restart files: f, p, ua, va, wa over the ghosted box
int64_t phased_state_doubles(const PhasedState *s);
int phased_get_state(PhasedState *s, double *host, std::string &err);
int phased_set_state(PhasedState *s, const double *host, std::string &err);

}  // namespace imx
