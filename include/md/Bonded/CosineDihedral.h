// Same header name as the reference's include/Bonded/CosineDihedral.h.
#include "Bonded/Terms.h"
