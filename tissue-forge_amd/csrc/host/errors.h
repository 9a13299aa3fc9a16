// Error convention of the C ABI: HRESULT-style codes (0 = S_OK, negative =
// failure; include/tf_port.h:255-282 of the reference) plus a thread-local
// message, the role of the reference's error stack (tfError.h:78).
#pragma once

#include <string>

#define TFG_OK 0
#define TFG_E_FAIL ((int)0x80004005)
#define TFG_E_INVALIDARG ((int)0x80070057)
#define TFG_E_OUTOFMEMORY ((int)0x8007000E)
#define TFG_E_RANGE ((int)0x8000000B)
#define TFG_E_STATE ((int)0x8000FFFF)
#define TFG_E_DEVICE ((int)0x80004006)
#define TFG_E_TOOFAST ((int)0x80004007)

namespace tfg {
int set_error(int code, const std::string &msg);
const char *last_error();
void clear_error();
}  // namespace tfg
