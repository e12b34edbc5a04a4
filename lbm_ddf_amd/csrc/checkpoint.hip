// checkpoint.hip -- the reference's binary dump formats and restart.
//
//   flow file  (output_flow, src/output.cu:67-146):  rho', u_x, u_y, u_z, phi,
//              each LX*LY*LZ doubles, x fastest, native (little-endian) order;
//   fg file    (output_fg, src/output.cu:379-419):   f then g, 27*LX*LY*LZ
//              doubles each, [ip][iz][iy][ix];
//   restart    (src/initial.cu:110-205): read both, then recompute the force
//              from phi exactly as the next step's forcing would
//              (src/initial.cu:202), so a restarted run continues bit for bit.
//
// Files hold FP64 values whatever the context's storage precision.
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

struct FileCloser {
    void operator()(FILE *f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

int io_fail(const std::string &what, const char *path) {
    return lbm::ctx_fail(LBM_E_IO, what + ": " + (path ? path : "(null)"));
}

const int32_t kFlowFields[5] = {LBM_FIELD_RHO, LBM_FIELD_UX, LBM_FIELD_UY, LBM_FIELD_UZ, LBM_FIELD_PHI};

}  // namespace

extern "C" {

int lbm_output_flow(lbm_ctx *c, const char *path) {
    if (!c || !path) return lbm::ctx_fail(LBM_E_INVALID, "NULL argument");
    std::vector<double> buf(c->n);
    File f(std::fopen(path, "wb"));
    if (!f) return io_fail("cannot open", path);
    for (int32_t field : kFlowFields) {
        if (int rc = lbm_get_field(c, field, buf.data())) return rc;
        if (std::fwrite(buf.data(), sizeof(double), c->n, f.get()) != c->n) return io_fail("short write", path);
    }
    if (std::fflush(f.get()) != 0) return io_fail("flush failed", path);
    return LBM_OK;
}

int lbm_output_fg(lbm_ctx *c, const char *path) {
    if (!c || !path) return lbm::ctx_fail(LBM_E_INVALID, "NULL argument");
    std::vector<double> buf(27 * c->n);
    File f(std::fopen(path, "wb"));
    if (!f) return io_fail("cannot open", path);
    for (int32_t which : {LBM_POP_F, LBM_POP_G}) {
        if (int rc = lbm_get_pops(c, which, buf.data())) return rc;
        if (std::fwrite(buf.data(), sizeof(double), buf.size(), f.get()) != buf.size())
            return io_fail("short write", path);
    }
    if (std::fflush(f.get()) != 0) return io_fail("flush failed", path);
    return LBM_OK;
}

int lbm_restart(lbm_ctx *c, const char *flow_path, const char *fg_path, int64_t step) {
    if (!c || !flow_path || !fg_path || step < 0) return lbm::ctx_fail(LBM_E_INVALID, "bad argument");
    File ff(std::fopen(flow_path, "rb"));
    if (!ff) return io_fail("cannot open", flow_path);
    File fg(std::fopen(fg_path, "rb"));
    if (!fg) return io_fail("cannot open", fg_path);
    // The state becomes defined before the uploads (set_* need a phase).
    if (c->phase == lbm::PHASE_EMPTY)
        if (int rc = lbm_init_arrays(c)) return rc;
    std::vector<double> buf(c->n);
    for (int32_t field : kFlowFields) {
        if (std::fread(buf.data(), sizeof(double), c->n, ff.get()) != c->n) return io_fail("short read", flow_path);
        if (int rc = lbm_set_field(c, field, buf.data())) return rc;
    }
    buf.resize(27 * c->n);
    for (int32_t which : {LBM_POP_F, LBM_POP_G}) {
        if (std::fread(buf.data(), sizeof(double), buf.size(), fg.get()) != buf.size())
            return io_fail("short read", fg_path);
        if (int rc = lbm_set_pops(c, which, buf.data())) return rc;
    }
    // forcing<<<>>> after the reads (src/initial.cu:202); precision-independent.
    if (int rc = lbm::ctx_forcing(c)) return rc;
    c->nstep = step;
    return LBM_OK;
}

}  // extern "C"
