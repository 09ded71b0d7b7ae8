/* GOSD.h -- the reference's include/GOSD.h name, served by the MI355X compat layer. */
#pragma once
#include "spm_reference_api.h"
