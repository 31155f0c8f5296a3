// ildg_abi.cpp -- C-ABI entry points for ILDG configuration files.
#include <string>
#include <vector>

#include "../../include/qcb.h"
#include "lime.hpp"

int qcb_set_free_error(int code, const std::string& msg);

extern "C" int qcb_ildg_write(const char* path, const int dims[4], const double* aos) {
    if (!path || !dims || !aos) return QCB_EINVAL;
    std::string err;
    if (!qcb_host::ildg_write(path, dims, aos, err)) return qcb_set_free_error(QCB_EINVAL, err);
    return QCB_OK;
}

extern "C" int qcb_ildg_read(const char* path, const int dims[4], double* aos) {
    if (!path || !dims || !aos) return QCB_EINVAL;
    std::string err;
    if (!qcb_host::ildg_read(path, dims, aos, err)) return qcb_set_free_error(QCB_EINVAL, err);
    return QCB_OK;
}

// Dimensions of a lattice handle, implemented beside the handle type.
int qcb_lattice_dims_internal(qcb_lattice* lat, int dims[4]);

extern "C" int qcb_ildg_save(qcb_lattice* lat, const char* path) {
    int d[4];
    if (!lat || !path || qcb_lattice_dims_internal(lat, d) != QCB_OK) return QCB_EINVAL;
    std::vector<double> h((size_t)d[0] * d[1] * d[2] * d[3] * 72);
    int rc = qcb_lattice_get_links(lat, h.data());
    if (rc) return rc;
    return qcb_ildg_write(path, d, h.data());
}

extern "C" int qcb_ildg_load(qcb_lattice* lat, const char* path) {
    int d[4];
    if (!lat || !path || qcb_lattice_dims_internal(lat, d) != QCB_OK) return QCB_EINVAL;
    std::vector<double> h((size_t)d[0] * d[1] * d[2] * d[3] * 72);
    int rc = qcb_ildg_read(path, d, h.data());
    if (rc) return rc;
    return qcb_lattice_set_links(lat, h.data());
}
