// CPU driver of the host loop's bounded polling wait (fsibe_2d_amd/csrc/
// wait.hpp) with an injected query and clock: no GPU and no HIP runtime.
// Prints one line per case: case rc elapsed_ms polls message
#include <cstdio>
#include <string>

#include "../../fsibe_2d_amd/csrc/wait.hpp"

using namespace fsibe;

int main() {
  // 1. never completes; a clock that advances 0.5 ms per reading
  {
    double t = 0.0;
    long polls = 0;
    std::string msg;
    const int rc = bounded_wait([&] { ++polls; return Poll::Busy; }, [&] { return t += 0.5; },
                                [] { return "unused"; }, 250.0, "the step's results", msg);
    std::printf("never %d %.1f %ld %s\n", rc, t, polls, msg.c_str());
  }
  // 2. completes at the 1000th poll, well before the deadline
  {
    double t = 0.0;
    long polls = 0;
    std::string msg;
    const int rc = bounded_wait([&] { return ++polls >= 1000 ? Poll::Done : Poll::Busy; },
                                [&] { return t += 0.5; }, [] { return "unused"; }, 250.0, "dt1", msg);
    std::printf("done %d %.1f %ld %s\n", rc, t, polls, msg.empty() ? "-" : msg.c_str());
  }
  // 3. the runtime reports an error at the 10th poll
  {
    double t = 0.0;
    long polls = 0;
    std::string msg;
    const int rc = bounded_wait([&] { return ++polls >= 10 ? Poll::Failed : Poll::Busy; },
                                [&] { return t += 0.5; }, [] { return "an illegal memory access"; }, 250.0,
                                "dt1", msg);
    std::printf("failed %d %.1f %ld %s\n", rc, t, polls, msg.c_str());
  }
  // 4. the deadline from the environment
  std::printf("timeout %.1f\n", wait_timeout_ms());
  return 0;
}
