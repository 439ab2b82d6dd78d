// search_i8.hip — instances of the search kernel for traversal on the crate's int8 codes (the crate's rule), with the exact f32 re-rank.
#include "search_impl.h"

HH_SEARCH_INST_CLASS(hhk::kPrecI8)
