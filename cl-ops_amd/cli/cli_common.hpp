// Command-line parsing shared by the bench programs: the GOption forms the
// reference's benchmarks accept (-x VALUE, -xVALUE, --long VALUE,
// --long=VALUE, flags without a value), with "Error: ..." on stderr and
// CLO_ERROR_ARGS on a bad option.

#pragma once

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cl_ops.h"

namespace cli {

enum Kind { kString, kUInt, kInt, kFlag };

struct Option {
	char short_name;
	const char* long_name;
	Kind kind;
	void* target;  // std::string*, unsigned long long*, long long* or bool*
	bool* seen;    // set when the option is given (may be NULL)
};

inline bool parse_number(const char* s, Kind kind, void* target) {
	char* end = nullptr;
	errno = 0;
	if (kind == kUInt) {
		if (*s == '-') return false;
		*static_cast<unsigned long long*>(target) = strtoull(s, &end, 0);
	} else {
		*static_cast<long long*>(target) = strtoll(s, &end, 0);
	}
	return errno == 0 && end && end != s && *end == '\0';
}

// Returns 0, or CLO_ERROR_ARGS after printing the error.  --help prints
// `usage` to stdout and exits with status 0.
inline int parse(int argc, char** argv, const std::vector<Option>& opts, const char* usage) {
	for (int i = 1; i < argc; ++i) {
		const char* arg = argv[i];
		if (!strcmp(arg, "--help") || !strcmp(arg, "-?")) {
			fputs(usage, stdout);
			exit(0);
		}
		const Option* opt = nullptr;
		const char* val = nullptr;
		if (arg[0] == '-' && arg[1] == '-') {
			const char* name = arg + 2;
			const char* eq = strchr(name, '=');
			const size_t len = eq ? static_cast<size_t>(eq - name) : strlen(name);
			for (const Option& o : opts)
				if (strlen(o.long_name) == len && !strncmp(o.long_name, name, len)) opt = &o;
			if (eq) val = eq + 1;
		} else if (arg[0] == '-' && arg[1]) {
			for (const Option& o : opts)
				if (o.short_name == arg[1]) opt = &o;
			if (arg[2]) val = arg + 2;
		}
		if (!opt) {
			fprintf(stderr, "Error: Unknown option %s\n", arg);
			return CLO_ERROR_ARGS;
		}
		if (opt->seen) *opt->seen = true;
		if (opt->kind == kFlag) {
			*static_cast<bool*>(opt->target) = true;
			continue;
		}
		if (!val) {
			if (i + 1 >= argc) {
				fprintf(stderr, "Error: Missing argument for %s\n", arg);
				return CLO_ERROR_ARGS;
			}
			val = argv[++i];
		}
		if (opt->kind == kString) {
			*static_cast<std::string*>(opt->target) = val;
		} else if (!parse_number(val, opt->kind, opt->target)) {
			fprintf(stderr, "Error: Cannot parse integer value '%s' for %s\n", val, arg);
			return CLO_ERROR_ARGS;
		}
	}
	return 0;
}

// Prints a library error as the reference's benchmarks do and returns its
// code as the exit status.
inline int fail(GError* err) {
	const int status = err ? err->code : CLO_ERROR_ARGS;
	fprintf(stderr, "Error: %s\n", err && err->message ? err->message : "unknown");
	clo_error_free(err);
	return status;
}

}  // namespace cli
