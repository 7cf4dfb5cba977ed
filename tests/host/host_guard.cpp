// Host build of gpluto_dev_amd/csrc/c_guard.hpp for tests/test_c_guard.py:
// a body that returns, and bodies that throw what host containers and
// std::thread throw, behind an extern "C" frame as in the library.
#include "../../gpluto_dev_amd/csrc/c_guard.hpp"

#include <cstdint>
#include <new>
#include <stdexcept>
#include <system_error>
#include <vector>

using namespace gpl;

extern "C" int hg_guard(int kind)
{
  return c_guard(-3, [&]() -> int {
    switch (kind) {
    case 1: throw std::bad_alloc();
    case 2: throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again));
    case 3: {
      std::vector<double> v;
      v.reserve(v.max_size() + 1);        // std::length_error, nothing allocated
      return (int)v.capacity();
    }
    case 4: throw 42;
    default: return 7;
    }
  });
}

extern "C" uint32_t hg_guard_u32(int kind)
{
  return c_guard<uint32_t>(0x80000000u, [&]() -> uint32_t {
    if (kind) throw std::runtime_error("step failed");
    return 5u;
  });
}
