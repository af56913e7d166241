// Temporary: entry points whose implementation files are not written yet.
#include "fip.h"
#include "fip_context.hpp"
using namespace fip;
extern "C" {
int fip_project(const double*, size_t, const double*, size_t, float*, float*) { return set_error(FIP_ERR_ARG, "not implemented"); }
int fip_collapse(const float*, const float*, const float*, const float*, const unsigned char*, const unsigned char*, const float*, const float*, size_t, float*, float*) { return set_error(FIP_ERR_ARG, "not implemented"); }
int fip_server_create(const fip_stream_config*, int, int, size_t, fip_server**) { return set_error(FIP_ERR_ARG, "not implemented"); }
int fip_server_submit(fip_server*, int, const void*, const void*, const float*, const float*, double) { return FIP_ERR_ARG; }
int fip_server_drain(fip_server*) { return FIP_ERR_ARG; }
int fip_server_stats(fip_server*, double*, double*, double*, long*, long*) { return FIP_ERR_ARG; }
int fip_server_latest(fip_server*, int, float*, long*) { return FIP_ERR_ARG; }
void fip_server_destroy(fip_server*) {}
int fip_fits_read(const char*, float**, long*) { return FIP_ERR_IO; }
int fip_fits_read_ext(const char*, const char*, float**, long*) { return FIP_ERR_IO; }
int fip_fits_write(const char*, const float*, long, long) { return FIP_ERR_IO; }
}
