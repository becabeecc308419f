// transformHIP.h -- the adapter a reference maintainer adds as src/transformHIP.h: the reference's
// abstract squaring engine (class transform, src/transform.h:42-256) implemented by this
// repository's C ABI (include/gfe.h, libgfe.so).  Compiled with the reference's own transform.h,
// gint.h and file.h in the include path; the factory branch that returns it from create_gpu is
// shown in INTEGRATION.md.  tests/test_adapter.py compiles and links it against the reference's
// header when the reference checkout is present.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>
#include "transform.h"
#include "gfe.h"                       // this repository's include/gfe.h

class transformHIP : public transform
{
private:
	gfe_engine * _e = nullptr;
	const size_t _num_regs;
	static void ck(const int rc) { if (rc != GFE_OK) throw std::runtime_error(std::string("gfe: ") + gfe_last_error()); }
	// the kind the reference's GPU engine reports for (b, n): NTT2 or NTT3 (src/transformGPU.h:947)
	static EKind kind_of(const uint32_t b, const uint32_t n)
	{ int rns = 3, is32 = 0; gfe_select(b, n, &rns, &is32); return (rns == 2) ? EKind::NTT2 : EKind::NTT3; }

public:
	transformHIP(const uint32_t b, const uint32_t n, const size_t device, const size_t num_regs)
		: transform(size_t(1) << n, n, b, kind_of(b, n)), _num_regs(num_regs)
	{ ck(gfe_create(b, n, uint32_t(num_regs), int(device), &_e)); }
	virtual ~transformHIP() { gfe_destroy(_e); }

	void set(const uint32_t a) override { ck(gfe_set(_e, a)); }
	void squareDup(const bool dup) override { ck(gfe_square_dup(_e, dup ? 1 : 0)); }
	void squareMul(const int32_t a) override { ck(gfe_square_mul(_e, a)); }
	void initMultiplicand(const size_t src) override { ck(gfe_init_multiplicand(_e, uint32_t(src))); }
	void mul() override { ck(gfe_mul(_e)); }
	void copy(const size_t dst, const size_t src) const override { ck(gfe_copy(_e, uint32_t(dst), uint32_t(src))); }
	size_t getMemSize() const override { return gfe_mem_size(_e); }
	size_t getCacheSize() const override { return gfe_cache_size(_e); }
	double getError() const override { return gfe_get_error(_e); }

	bool readContext(file & cFile, const size_t nregs) override
	{
		const uint32_t nr = uint32_t((nregs != 0) ? nregs : _num_regs);
		std::vector<char> buf(gfe_context_size(_e, nr));
		if (!cFile.read(buf.data(), buf.size())) return false;
		return gfe_load(_e, nr, buf.data(), buf.size()) == GFE_OK;
	}
	void saveContext(file & cFile, const size_t nregs) const override
	{
		const uint32_t nr = uint32_t((nregs != 0) ? nregs : _num_regs);
		std::vector<char> buf(gfe_context_size(_e, nr));
		ck(gfe_save(_e, nr, buf.data(), buf.size()));
		cFile.write(buf.data(), buf.size());
	}

protected:
	void getZi(int32_t * const zi) const override { ck(gfe_get_digits(_e, 0, zi)); }
	void setZi(const int32_t * const zi) override { ck(gfe_set_digits(_e, 0, zi)); }
};
