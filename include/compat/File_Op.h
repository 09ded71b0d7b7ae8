/* File_Op.h -- the reference's include/File_Op.h name, served by the MI355X compat layer. */
#pragma once
#include "spm_reference_api.h"
