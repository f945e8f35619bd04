// Same header name as the reference's include/Bonded/harmonicUB.h.
#include "Bonded/Terms.h"
