// search_any.hip — instances of the search kernel for mixed batches: every precision, chosen per query.
#include "search_impl.h"

HH_SEARCH_INST_CLASS(hhk::kPrecAny)
