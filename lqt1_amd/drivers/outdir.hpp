// Output folders of the driver programs: created when missing (the
// reference programs expect Outputs/, Outputs2Dreal/ and Output3D/ to exist).
#pragma once

#include <errno.h>
#include <sys/stat.h>

#include <string>

// mkdir -p: every component of `dir`; true when the folder exists afterwards.
inline bool make_dirs(const std::string& dir) {
    if (dir.empty()) return true;
    for (size_t i = 1; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            const std::string part = dir.substr(0, i);
            if (mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return false;
        }
    }
    struct stat st;
    return stat(dir.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// The folder part of a file path ("" for a bare file name).
inline std::string parent_dir(const std::string& file) {
    const size_t k = file.rfind('/');
    return k == std::string::npos ? std::string() : file.substr(0, k);
}
