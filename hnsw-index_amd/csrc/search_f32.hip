// search_f32.hip — instances of the search kernel for uniform f32 batches
// (141 VGPRs without the other precisions' paths, against 168 in the
// every-precision instance).
#include "search_impl.h"

HH_SEARCH_INST_CLASS(hhk::kPrecF32)
