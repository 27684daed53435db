// Last-error plumbing of the C-ABI: C++ exceptions and HIP failures are
// turned into integer status codes plus a thread-local message.
#pragma once
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/psipm_hip.h"

namespace psg {

void set_error(const std::string& msg);

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return PSG_ERR_MEMORY;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PSG_ERR_INTERNAL;
  } catch (...) {
    set_error("unknown exception");
    return PSG_ERR_INTERNAL;
  }
}

}  // namespace psg
