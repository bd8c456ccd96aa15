// CenterlineExtraction.h (3D) -- correctly spelt alias of CenterlineExctraction.h.
#pragma once

#include "CenterlineExctraction.h"
