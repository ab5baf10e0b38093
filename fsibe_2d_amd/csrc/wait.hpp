// wait.hpp -- bounded polling waits of the host loop.
//
// The time loop waits for a step's scalars (dt1) or results by polling an
// event, so that the waiting thread stays on its core and the host work that
// follows starts at once.  A poll without a deadline turns a kernel that
// never finishes into a silent busy loop; these waits give up after a
// deadline and report which stage they were waiting for.  The header has no
// HIP types: the query and the clock are passed in, so the logic also builds
// in a plain C++ program.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/fsibe_hip.h"

namespace fsibe {

// one poll of the awaited work
enum class Poll { Done, Busy, Failed };

// Deadline of a polling wait: FSIBE_WAIT_TIMEOUT_MS milliseconds if set to a
// positive value, else 120000.  A step of the largest grids takes a few
// milliseconds, so only work that never completes reaches it.
inline double wait_timeout_ms() {
  const char *v = std::getenv("FSIBE_WAIT_TIMEOUT_MS");
  if (v && *v) {
    const double t = std::atof(v);
    if (t > 0.0) return t;
  }
  return 120000.0;
}

// Poll query() (returning Poll) until it is Done or Failed, reading the clock
// now_ms() every 64 polls; after timeout_ms without completion, returns
// FSIBE_EHIP with `msg` naming `stage`.  On Failed, `msg` names the stage and
// describe() (a const char *).
template <class Query, class Now, class Describe>
int bounded_wait(Query &&query, Now &&now_ms, Describe &&describe, double timeout_ms, const char *stage,
                 std::string &msg) {
  const double t0 = now_ms();
  for (unsigned k = 1;; ++k) {
    const Poll r = query();
    if (r == Poll::Done) return FSIBE_OK;
    if (r == Poll::Failed) {
      char b[384];
      std::snprintf(b, sizeof b, "waiting for %s: %s", stage, describe());
      msg = b;
      return FSIBE_EHIP;
    }
    if ((k & 63u) == 0u) {
      const double el = now_ms() - t0;
      if (el > timeout_ms) {
        char b[384];
        std::snprintf(b, sizeof b,
                      "waiting for %s: not complete after %.0f ms (deadline FSIBE_WAIT_TIMEOUT_MS = %.0f ms); "
                      "the handle's stream may hold a kernel that does not finish",
                      stage, el, timeout_ms);
        msg = b;
        return FSIBE_EHIP;
      }
    }
  }
}

}  // namespace fsibe
