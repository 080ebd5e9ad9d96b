// Compile-time column transforms: instances of the lengths in ASARF_COL_LENGTHS_2 (col_static.h).
#include "col_static.h"

namespace asarf {

bool launch_col_static_2(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err) {
    switch (a.sub.n) {
        ASARF_COL_LENGTHS_2(ASARF_COL_LAUNCH)
        default: return false;
    }
}

}  // namespace asarf
