// Same header name as the reference's include/Bonded/HarmonicImproper.h.
#include "Bonded/Terms.h"
