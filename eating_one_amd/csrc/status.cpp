// Library identity and error reporting: ABI version, source hash, status
// strings and the per-thread message of the last failure.  Host-only code.
#include <string>

#include "ctx.hpp"
#include "eojoin.h"

namespace eo {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

}  // namespace eo

using namespace eo;

extern "C" {

uint32_t eo_abi_version(void) { return EOJOIN_ABI_VERSION; }

#ifndef EO_SOURCE_HASH
#define EO_SOURCE_HASH "unknown"
#endif
const char *eo_build_source_hash(void) { return EO_SOURCE_HASH; }

const char *eo_status_string(eo_status st) {
  switch (st) {
    case EO_OK: return "ok";
    case EO_EINVAL: return "invalid argument";
    case EO_ENOMEM: return "out of memory";
    case EO_ECAPACITY: return "output capacity too small";
    case EO_EHIP: return "HIP runtime error";
    case EO_EBUDGET: return "HBM budget exceeded";
    case EO_ENODEV: return "no usable GPU";
    case EO_EAGAIN: return "not ready";
    case EO_ECANCELED: return "cancelled";
    case EO_ETIMEDOUT: return "deadline expired";
    case EO_EINTERNAL: return "internal error";
  }
  return "unknown status";
}

const char *eo_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
