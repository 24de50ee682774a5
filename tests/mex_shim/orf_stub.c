/* Stand-in for liborf_rcwa.so on machines without a GPU (test scaffolding):
 * orf_rcwa2d reports progress twice and returns the status in ORF_STUB_STATUS,
 * so the MEX gateway's error mapping and cancel path run on the CPU. */
#include <stdio.h>
#include <stdlib.h>

#include "orf_rcwa.h"

struct orf_ctx { int device; };
static struct orf_ctx g_stub_ctx;
static char g_msg[256] = "";

int orf_init(int device, orf_ctx** ctx) {
  g_stub_ctx.device = device;
  *ctx = &g_stub_ctx;
  return ORF_OK;
}
void orf_release(orf_ctx* ctx) { (void)ctx; }
const char* orf_last_error(void) { return g_msg; }
int orf_harmonic_count(const orf_problem* p) { return (2 * p->Nx + 1) * (2 * p->Ny + 1); }
int orf_rcwa2d(orf_ctx* ctx, const orf_problem* p, orf_result* out, orf_progress_fn progress, void* user) {
  (void)ctx;
  const char* s = getenv("ORF_STUB_STATUS");
  const int status = s ? atoi(s) : ORF_OK;
  if (progress) {
    if (progress(0.5, "layer 1/1", user) || progress(1.0, "done", user)) {
      snprintf(g_msg, sizeof g_msg, "cancelled by the progress callback");
      return ORF_CANCELLED;
    }
  }
  if (status != ORF_OK) {
    snprintf(g_msg, sizeof g_msg, "stub status %d", status);
    return status;
  }
  const int M = orf_harmonic_count(p);
  const long nout = (long)p->nStruct * p->nLambda * p->nPol;
  for (long k = 0; k < M * nout; ++k) out->R[k] = 0.25;
  return ORF_OK;
}
