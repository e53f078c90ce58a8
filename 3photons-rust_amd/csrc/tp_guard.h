// tp_guard.h -- no C++ exception leaves the C ABI.  Every extern "C" entry
// point runs its body through tp_guard: a failed host allocation becomes
// TP_ERR_NOMEM, anything else TP_ERR_INTERNAL, with the message stored where
// tp_last_error finds it when the call has a context.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "tp.h"

template <class F> int tp_guard(std::string *err, F &&body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        try {
            if (err)
                *err = "host allocation failed";
        } catch (...) {
        }
        return TP_ERR_NOMEM;
    } catch (const std::exception &e) {
        try {
            if (err)
                *err = std::string("internal error: ") + e.what();
        } catch (...) {
        }
        return TP_ERR_INTERNAL;
    } catch (...) {
        try {
            if (err)
                *err = "internal error";
        } catch (...) {
        }
        return TP_ERR_INTERNAL;
    }
}
