// hsba_server.cpp — continuous-batching serving loop (placeholder until the
// scheduler lands): every entry point reports HSBA_ERR_UNSUPPORTED.
#include "hsba_host.h"

extern "C" {
struct hsba_server { int unused; };
int hsba_server_create(const hsba_server_options *, hsba_server **server) {
  if (server) *server = nullptr;
  return HSBA_ERR_UNSUPPORTED;
}
int hsba_server_destroy(hsba_server *server) { delete server; return HSBA_OK; }
int hsba_server_submit(hsba_server *, const hsba_problem_desc *, const double *, const double *,
                       const hsba_ycov_inverse *, double *, int64_t *) { return HSBA_ERR_UNSUPPORTED; }
int hsba_server_poll(hsba_server *, hsba_job_result *, int32_t, int32_t *count) {
  if (count) *count = 0;
  return HSBA_ERR_UNSUPPORTED;
}
int hsba_server_drain(hsba_server *) { return HSBA_ERR_UNSUPPORTED; }
double hsba_server_clock(const hsba_server *) { return 0.0; }
}
