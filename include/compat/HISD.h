/* HISD.h -- the reference's include/HISD.h name, served by the MI355X compat layer. */
#pragma once
#include "spm_reference_api.h"
