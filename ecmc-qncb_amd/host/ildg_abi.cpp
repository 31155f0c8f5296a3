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

// qcb_ildg_save / qcb_ildg_load stage through the device: csrc/qcb_io.cpp.
