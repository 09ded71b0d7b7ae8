/* My_FFT.h -- the reference's include/My_FFT.h name, served by the MI355X compat layer. */
#pragma once
#include "spm_reference_api.h"
