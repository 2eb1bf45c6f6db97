// "-name value" command-line options with the conventions of the reference's
// tools (src/utils/cmd_option.h, used by src/main_vnlb.cpp:37-86):
//   * an option is looked up by exact name anywhere in argv; its value is
//     the next word, or the name itself when it is the last word;
//   * -h, -help or --help anywhere lists every option queried, with its
//     value, on stderr, and the program continues;
//   * booleans are true unless the value is false/off/0 (any case);
//     integers parse with atoi; floats accept a fraction "a/b".
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <strings.h>

class CliOptions {
public:
	CliOptions(int argc, char **argv, const char *title, const char *note = nullptr) : argc_(argc), argv_(argv)
	{
		help_ = find("-h") || find("-help") || find("--help");
		if (help_) {
			const char *slash = std::strrchr(argv[0], '/');
			std::fprintf(stderr, "\n %s : %s\n\n", slash ? slash + 1 : argv[0], title);
			if (note) std::fprintf(stderr, "%s\n", note);
		}
	}

	std::string get(const char *name, const char *def, const char *usage)
	{
		const char *s = find(name);
		const std::string v = s ? s : def;
		show(name, v.empty() ? "0" : v.c_str(), usage);
		return v;
	}
	bool get(const char *name, bool def, const char *usage)
	{
		const char *s = find(name);
		const bool v = s ? !(strcasecmp(s, "false") == 0 || strcasecmp(s, "off") == 0 || strcasecmp(s, "0") == 0) : def;
		show(name, v ? "true" : "false", usage);
		return v;
	}
	int get(const char *name, int def, const char *usage)
	{
		const char *s = find(name);
		const int v = s ? std::atoi(s) : def;
		show(name, std::to_string(v).c_str(), usage);
		return v;
	}
	float get(const char *name, float def, const char *usage)
	{
		const char *s = find(name);
		float v = def;
		if (s) {
			float num = 0.f, den = 1.f;
			std::sscanf(s, "%g/%g", &num, &den);
			v = num / den;
		}
		char buf[64];
		std::snprintf(buf, sizeof buf, "%g", v);
		show(name, buf, usage);
		return v;
	}

private:
	const char *find(const char *name) const
	{
		for (int k = 0; k < argc_; k++)
			if (std::strcmp(argv_[k], name) == 0) return k + 1 < argc_ ? argv_[k + 1] : argv_[k];
		return nullptr;
	}
	void show(const char *name, const char *value, const char *usage) const
	{
		if (help_) std::fprintf(stderr, "    %-8s = %-12s : %s\n", name, value, usage);
	}

	int argc_;
	char **argv_;
	bool help_ = false;
};
