// Shared host helpers: error state, status exceptions.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/baandi_amd.h"

namespace bd {

// An error carrying the C status it maps to.
struct Error : std::runtime_error {
  bd_status status;
  Error(bd_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& msg);
void clear_last_error();

[[noreturn]] inline void fail(bd_status s, const std::string& m) { throw Error(s, m); }

}  // namespace bd

// Wrap a C-ABI body: exceptions become status codes plus the message.
#define BD_GUARD(...)                                                \
  try {                                                              \
    bd::clear_last_error();                                          \
    __VA_ARGS__;                                                     \
    return BD_OK;                                                    \
  } catch (const bd::Error& e) {                                     \
    bd::set_last_error(e.what());                                    \
    return e.status;                                                 \
  } catch (const std::exception& e) {                                \
    bd::set_last_error(e.what());                                    \
    return BD_ERR_APPLICATION;                                       \
  }
