/* Internal hooks shared by the library's host files (not part of the C ABI). */
#ifndef LRH_INTERNAL_H
#define LRH_INTERNAL_H

#include <string>

/* Sets the calling thread's lrh_last_error() message and returns `code`. */
int lrhSetError(int code, const std::string &msg);

#endif
