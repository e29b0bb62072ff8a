// Status codes and the per-thread error text of the C-ABI: every entry point runs its body
// through guarded(), which maps the library's exception types to qh_status values and keeps
// the message for qh_hip_last_error().
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/qoraholo_hip.h"
#include "runtime.hpp"

namespace qh {

// Message of the last failed call on the calling thread.
std::string& last_error();

template <class F>
int guarded(F&& f) {
  try {
    f();
    last_error().clear();
    return QH_OK;
  } catch (const InvalidArgument& e) {
    last_error() = e.what();
    return QH_ERR_INVALID_ARGUMENT;
  } catch (const LogicError& e) {
    last_error() = e.what();
    return QH_ERR_LOGIC;
  } catch (const OutOfMemory& e) {
    last_error() = e.what();
    return QH_ERR_OUT_OF_MEMORY;
  } catch (const std::bad_alloc&) {
    last_error() = "out of host memory";
    return QH_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return QH_ERR_RUNTIME;
  }
}

}  // namespace qh
