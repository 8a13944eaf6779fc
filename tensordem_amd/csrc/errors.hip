// Error string storage for the C ABI.
#include "common.h"

namespace tdm {

std::string &last_error()
{
  static thread_local std::string msg;
  return msg;
}

int fail(int code, const std::string &msg)
{
  last_error() = msg;
  return code;
}

}  // namespace tdm
