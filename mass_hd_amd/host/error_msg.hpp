// error_msg.hpp -- error and warning exceptions of the host program.
//
// Same contract as the reference's inline_msg (src/error_msg.hpp:12-56): the
// constructor streams all its arguments into one message; error_msg always
// stops the run, warning_msg only with --stop-at-warning.
#pragma once

#include <exception>
#include <sstream>
#include <string>
#include <utility>

template <typename Tag>
class inline_msg : public std::exception {
  std::string msg_;

 public:
  template <class... Args>
  explicit inline_msg(Args&&... args) {
    std::ostringstream s;
    (void)std::initializer_list<int>{((s << std::forward<Args>(args)), 0)...};
    msg_ = s.str();
  }
  const char* what() const noexcept override { return msg_.c_str(); }
};

using error_msg = inline_msg<struct ErrorTag>;
using warning_msg = inline_msg<struct WarningTag>;

template <class... Args>
std::string inline_str(Args&&... args) {
  std::ostringstream s;
  (void)std::initializer_list<int>{((s << std::forward<Args>(args)), 0)...};
  return s.str();
}
