// tools::WriteNrg / tools::SaveNrg -- the reference's src/io_tools/write_nrg.h:65-75, declared with
// the reader in mbx/read_nrg.h (one translation unit holds both); this header keeps the
// reference's include name for callers that include the writer alone.
#pragma once

#include "mbx/read_nrg.h"
