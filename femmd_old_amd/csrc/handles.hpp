// handles.hpp -- the opaque C handles of include/femmd.h.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"

struct fm_system {
    fm::HostSystem hs;
    std::unique_ptr<fm::Engine> eng;
    std::string err;
};

struct fm_batch {
    std::vector<fm::HostSystem> hs;
    std::unique_ptr<fm::Engine> eng;
    std::string err;
};
