// common.hpp -- host helpers shared by the library's translation units.
#pragma once
#include <string>

#include "tsmi_hip.h"

namespace tsmi {

// Records msg as the calling thread's last error (tsmi_last_error) and returns code.
int set_error(int code, const std::string &msg);

// Validates a descriptor; `who` prefixes the message.  TSMI_OK or an error code.
int check_desc(const tsmi_desc &d, const char *who);

// Largest even segment height whose rows plus the apron rows of a pass
// (apron_rows in all: 8 for one step per pass, 16 for two) keep a wave's 32-bit
// byte offsets valid for a row pitch of `pitch` elements (FP64 planes): the
// buffer range of a segment stays below 2^31 bytes, so the offset 0x80000000
// that drops a store or zeroes a load always lies outside it.
int max_seg_rows(long long pitch, int apron_rows);

// Rows per segment of a pass over H rows cut into `strips` strips: tile_rows
// when > 0; otherwise one round of resident waves (resident = waves the device
// holds at the kernel's occupancy, 0 if unknown: half the rows), at least
// floor_rows -- or at least 12 where floor_rows-row segments would leave some of
// the `simds` SIMDs without a wave (small12) -- then clamped by max_seg_rows and
// rounded up to even (segments start on even rows).  Not above H.
int plan_seg_rows(int H, int strips, long resident, long simds, int floor_rows, bool small12, int tile_rows,
                  long long pitch, int apron_rows);

}  // namespace tsmi
