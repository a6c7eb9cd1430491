// nr_error.h — the per-thread "last error" message behind nr_last_error(),
// shared by the translation units of libnectar_hip.so.
#pragma once

#include <string>

std::string& nr_error_slot();
