// Exception-to-status wrapper of every C ABI entry point: 0 on success,
// 1 for a rejected call, 2 for a HIP error, 3 otherwise; the message is kept
// for hit3d_last_error.
#pragma once
#include <stdexcept>
#include <string>

#include "state.hpp"

namespace hit3d {

template <typename Fn>
inline int guarded(Fn&& fn, bool device = true) {
    try {
        fn();
        const hipError_t e = device ? hipGetLastError() : hipSuccess;
        if (e != hipSuccess) throw HipError{std::string("kernel launch: ") + hipGetErrorString(e)};
        return 0;
    } catch (const HipError& e) {
        lib().err = e.msg;
        return 2;
    } catch (const std::exception& e) {
        lib().err = e.what();
        return 1;
    } catch (...) {
        lib().err = "unknown error";
        return 3;
    }
}

}  // namespace hit3d
