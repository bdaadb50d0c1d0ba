// ensemble.hip — batched ensembles (gnss_ensemble_*).
#include <string>

#include "engine.h"

#define GNSS_API extern "C" __attribute__((visibility("default")))

struct gnss_ens {
  std::string err;
};

static thread_local std::string g_ens_err = "ensembles are not available in this build";

GNSS_API int gnss_ensemble_create(int, const gnss_params*, gnss_ens**) { return GNSS_ERR_UNSUPPORTED; }
GNSS_API int gnss_ensemble_destroy(gnss_ens*) { return GNSS_OK; }
GNSS_API const char* gnss_ensemble_last_error(const gnss_ens* e) { return e ? e->err.c_str() : g_ens_err.c_str(); }
GNSS_API int gnss_ensemble_init(gnss_ens*) { return GNSS_ERR_UNSUPPORTED; }
GNSS_API int gnss_ensemble_step(gnss_ens*, int) { return GNSS_ERR_UNSUPPORTED; }
GNSS_API int gnss_ensemble_stats(gnss_ens*, int, gnss_stats_out*, double*, double*) { return GNSS_ERR_UNSUPPORTED; }
GNSS_API int gnss_ensemble_get_field(gnss_ens*, int, int, int, double*) { return GNSS_ERR_UNSUPPORTED; }
