/* LPsrc.h -- the reference's include/LPsrc.h name, served by the MI355X compat layer. */
#pragma once
#include "spm_reference_api.h"
