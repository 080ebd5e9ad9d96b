// Dispatch of the compile-time column transforms (col_static.h) and the first unit of their instances.
#include "col_static.h"

namespace asarf {

// Sub-transform lengths with a compile-time kernel (the planner accepts them even where no runtime radix plan
// fits the 16 lanes of a column).
bool col_static_length(int n) {
    switch (n) {
        ASARF_COL_LENGTHS_0(ASARF_COL_LABEL)
        ASARF_COL_LENGTHS_1(ASARF_COL_LABEL)
        ASARF_COL_LENGTHS_2(ASARF_COL_LABEL)
        ASARF_COL_LENGTHS_3(ASARF_COL_LABEL)
        ASARF_COL_LENGTHS_4(ASARF_COL_LABEL)
        ASARF_COL_LENGTHS_5(ASARF_COL_LABEL)
            return true;
        default:
            return false;
    }
}

bool launch_col_static_0(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err) {
    switch (a.sub.n) {
        ASARF_COL_LENGTHS_0(ASARF_COL_LAUNCH)
        default: return false;
    }
}

// Returns true (and the launch status in *err) when a compile-time kernel exists for the sub-transform length.
bool launch_col_static(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err) {
    return launch_col_static_0(a, inverse, s, err) || launch_col_static_1(a, inverse, s, err) ||
           launch_col_static_2(a, inverse, s, err) || launch_col_static_3(a, inverse, s, err) ||
           launch_col_static_4(a, inverse, s, err) || launch_col_static_5(a, inverse, s, err);
}

}  // namespace asarf
