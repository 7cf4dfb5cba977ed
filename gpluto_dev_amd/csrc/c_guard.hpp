// The C ABI reports failure through return codes.  Host containers
// (std::vector, std::string) and std::thread report exhaustion by throwing
// (std::bad_alloc, std::length_error, std::system_error), and an exception
// that unwinds out of an extern "C" function ends the caller's process.  Every
// entry point that allocates on the host runs its body through c_guard, which
// turns any exception into the given error value.
#pragma once

namespace gpl {

template <typename R, typename F>
R c_guard(R on_error, F &&body) noexcept
{
  try {
    return body();
  } catch (...) {
    return on_error;
  }
}

}  // namespace gpl
