// placeholder: filled in below
#include "../../include/ffvhc.h"
