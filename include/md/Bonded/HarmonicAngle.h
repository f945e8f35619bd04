// Same header name as the reference's include/Bonded/HarmonicAngle.h.
#include "Bonded/Terms.h"
