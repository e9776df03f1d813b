/* Extended operators of the oracle (filled in below). */
#include "sfk_oracle.h"
