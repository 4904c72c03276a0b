// placeholder, replaced by batch.hip / server.hip
#include "../../include/eithot.h"
#include "runtime.hpp"
using namespace eithot;
extern "C" {
eithot_batch* eithot_batch_create(void) { return nullptr; }
int eithot_batch_add(eithot_batch*, void*, const int64_t*, const int*, int, size_t) { return EITHOT_ERR_UNSUPPORTED; }
int eithot_batch_build(eithot_batch*, double, int) { return EITHOT_ERR_UNSUPPORTED; }
int eithot_batch_run(eithot_batch*, void*, int) { return EITHOT_ERR_UNSUPPORTED; }
int eithot_batch_info(const eithot_batch*, int*, int64_t*, int*) { return EITHOT_ERR_UNSUPPORTED; }
void eithot_batch_destroy(eithot_batch*) {}
eithot_server* eithot_server_create(int64_t, int, int) { return nullptr; }
int64_t eithot_server_submit(eithot_server*, void*, const int64_t*, const int*, int, size_t) { return EITHOT_ERR_UNSUPPORTED; }
int eithot_server_poll(eithot_server*, int) { return EITHOT_ERR_UNSUPPORTED; }
int eithot_server_status(eithot_server*, int64_t, double*) { return EITHOT_ERR_UNSUPPORTED; }
int eithot_server_drain(eithot_server*) { return EITHOT_ERR_UNSUPPORTED; }
void eithot_server_destroy(eithot_server*) {}
}
