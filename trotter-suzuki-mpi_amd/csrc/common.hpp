// common.hpp -- host helpers shared by the library's translation units.
#pragma once
#include <string>

#include "tsmi_hip.h"

namespace tsmi {

// Records msg as the calling thread's last error (tsmi_last_error) and returns code.
int set_error(int code, const std::string &msg);

// Validates a descriptor; `who` prefixes the message.  TSMI_OK or an error code.
int check_desc(const tsmi_desc &d, const char *who);

// Largest even segment height whose rows plus the 4-row aprons keep a wave's
// 32-bit byte offsets valid for a row pitch of `pitch` elements.
int max_seg_rows(long long pitch);

}  // namespace tsmi
