// search_codes.hip — instances of the search kernel for traversal on the per-dimension int8 codes, with the exact f32 re-rank.
#include "search_impl.h"

HH_SEARCH_INST_CLASS(hhk::kPrecCodes)
