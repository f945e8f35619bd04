// Same header name as the reference's include/Bonded/CMapGroup.h.
#include "Bonded/Terms.h"
