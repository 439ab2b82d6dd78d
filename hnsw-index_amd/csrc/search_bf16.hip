// search_bf16.hip — instances of the search kernel for bf16 traversal with the exact f32 re-rank.
#include "search_impl.h"

HH_SEARCH_INST_CLASS(hhk::kPrecBf16)
