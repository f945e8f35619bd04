// Same header name as the reference's include/Bonded/HarmonicBond.h.
#include "Bonded/Terms.h"
