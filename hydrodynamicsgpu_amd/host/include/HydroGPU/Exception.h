// HydroGPU/Exception.h -- the error type of the host interface.
// The reference throws Common::Exception (a std::exception with a stream
// operator for the message); every failing C-ABI status is rethrown as this.
#pragma once
#include <sstream>
#include <stdexcept>
#include <string>

namespace HydroGPU {

struct Exception : public std::runtime_error {
  Exception() : std::runtime_error(""), msg_() {}
  explicit Exception(const std::string& m) : std::runtime_error(m), msg_(m) {}
  Exception(const Exception& o) : std::runtime_error(o.msg_), msg_(o.msg_) {}
  template <typename T>
  Exception& operator<<(const T& v) {
    std::ostringstream ss;
    ss << msg_ << v;
    msg_ = ss.str();
    return *this;
  }
  const char* what() const noexcept override { return msg_.c_str(); }

 private:
  std::string msg_;
};

// throws Exception(where + ": " + hydro_last_error()) when status != 0
void checkStatus(int status, const char* where);

}  // namespace HydroGPU
