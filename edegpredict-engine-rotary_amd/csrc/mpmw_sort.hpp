// mpmw_sort.hpp -- particle binning state (cell-ordered particle layout).
#pragma once

#include <stdint.h>

struct SortState {
    bool valid = false;        // particles are ordered by the current grid's bins
};
