// rbx_pending.hip -- entry points whose kernels are not built yet.  Each
// returns RBX_UNSUPPORTED with a message, so callers fail loudly.
#include "rbx_internal.hh"

using namespace rbx;

#define PENDING(name)                                                          \
    ::rbx::set_error("%s: not implemented in this build", name);               \
    return RBX_UNSUPPORTED;

extern "C" {

rbx_status rbx_sparse_skop_init(const rbx_sparse_dist *, const rbx_rng_state *, rbx_sparse_skop *) {
    PENDING("rbx_sparse_skop_init")
}
rbx_status rbx_sparse_dist_props(const rbx_sparse_dist *, int64_t *, int64_t *, int64_t *, double *) {
    PENDING("rbx_sparse_dist_props")
}
rbx_status rbx_sfill_sparse_unpacked_nosub(const rbx_sparse_dist *, int64_t *, float *, int64_t *, int64_t *,
                                           const rbx_rng_state *, rbx_rng_state *, void *) {
    PENDING("rbx_sfill_sparse_unpacked_nosub")
}
rbx_status rbx_dfill_sparse_unpacked_nosub(const rbx_sparse_dist *, int64_t *, double *, int64_t *, int64_t *,
                                           const rbx_rng_state *, rbx_rng_state *, void *) {
    PENDING("rbx_dfill_sparse_unpacked_nosub")
}
rbx_status rbx_sfill_sparse_skop(rbx_sparse_skop *, void *) { PENDING("rbx_sfill_sparse_skop") }
rbx_status rbx_dfill_sparse_skop(rbx_sparse_skop *, void *) { PENDING("rbx_dfill_sparse_skop") }
rbx_status rbx_sparse_skop_free(rbx_sparse_skop *) { PENDING("rbx_sparse_skop_free") }
rbx_status rbx_slskges(char, char, char, int64_t, int64_t, int64_t, float, const rbx_sparse_skop *, int64_t,
                       int64_t, const float *, int64_t, float, float *, int64_t, void *) {
    PENDING("rbx_slskges")
}
rbx_status rbx_dlskges(char, char, char, int64_t, int64_t, int64_t, double, const rbx_sparse_skop *, int64_t,
                       int64_t, const double *, int64_t, double, double *, int64_t, void *) {
    PENDING("rbx_dlskges")
}
rbx_status rbx_srskges(char, char, char, int64_t, int64_t, int64_t, float, const float *, int64_t,
                       const rbx_sparse_skop *, int64_t, int64_t, float, float *, int64_t, void *) {
    PENDING("rbx_srskges")
}
rbx_status rbx_drskges(char, char, char, int64_t, int64_t, int64_t, double, const double *, int64_t,
                       const rbx_sparse_skop *, int64_t, int64_t, double, double *, int64_t, void *) {
    PENDING("rbx_drskges")
}
rbx_status rbx_srsksp3(char, char, char, int64_t, int64_t, int64_t, float, const rbx_csr_matrix *, int64_t,
                       int64_t, const rbx_dense_skop *, int64_t, int64_t, float, float *, int64_t, void *) {
    PENDING("rbx_srsksp3")
}
rbx_status rbx_drsksp3(char, char, char, int64_t, int64_t, int64_t, double, const rbx_csr_matrix *, int64_t,
                       int64_t, const rbx_dense_skop *, int64_t, int64_t, double, double *, int64_t, void *) {
    PENDING("rbx_drsksp3")
}

}  // extern "C"
