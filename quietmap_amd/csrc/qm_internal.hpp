// The library's private handle types, shared by its translation units.
#pragma once

#include <mutex>

#include "qm_device.hpp"

struct qm_gpu {
    int ordinal;
    char name[256];
    qm::BandTable bands;
    std::mutex lock;
};
