// AlignedBuffer lives with the field containers (core/fields.hpp); this header keeps the
// reference's include path (core/aligned_buffer.hpp) valid.
#pragma once

#include "core/fields.hpp"
