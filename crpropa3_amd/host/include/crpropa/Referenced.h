// Intrusive reference counting with the reference's ownership model
// (include/crpropa/Referenced.h:27-215): objects derive from Referenced and
// are held by ref_ptr<T>; the last ref_ptr deletes the object.
#ifndef CRPROPA_AMD_REFERENCED_H
#define CRPROPA_AMD_REFERENCED_H

#include <atomic>
#include <cstddef>

namespace crpropa {

class Referenced {
public:
	Referenced() : refcount_(0) {}
	Referenced(const Referenced &) : refcount_(0) {}
	Referenced &operator=(const Referenced &) { return *this; }
	virtual ~Referenced() {}

	size_t addReference() const { return ++refcount_; }
	size_t removeReference() const {
		size_t left = --refcount_;
		if (left == 0)
			delete this;
		return left;
	}
	size_t getReferenceCount() const { return refcount_; }

private:
	mutable std::atomic<size_t> refcount_;
};

template <class T>
class ref_ptr {
public:
	ref_ptr() : p_(nullptr) {}
	ref_ptr(T *p) : p_(p) { if (p_) p_->addReference(); }
	ref_ptr(const ref_ptr &r) : p_(r.p_) { if (p_) p_->addReference(); }
	template <class U>
	ref_ptr(const ref_ptr<U> &r) : p_(r.get()) { if (p_) p_->addReference(); }
	~ref_ptr() { if (p_) p_->removeReference(); }

	ref_ptr &operator=(const ref_ptr &r) { assign(r.p_); return *this; }
	template <class U>
	ref_ptr &operator=(const ref_ptr<U> &r) { assign(r.get()); return *this; }
	ref_ptr &operator=(T *p) { assign(p); return *this; }

	T *get() const { return p_; }
	T &operator*() const { return *p_; }
	T *operator->() const { return p_; }
	operator T *() const { return p_; }
	bool valid() const { return p_ != nullptr; }
	bool operator!() const { return p_ == nullptr; }

private:
	void assign(T *p) {
		if (p == p_)
			return;
		T *old = p_;
		p_ = p;
		if (p_) p_->addReference();
		if (old) old->removeReference();
	}
	T *p_;
};

} // namespace crpropa

#endif
